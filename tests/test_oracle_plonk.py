"""The PLONK quotient oracle, pinned without the reference: batch inversion
against pow(x, -1); the NTT plumbing against Horner evaluation at every coset
point and division by x^n - 1; the numerator formula by valid PLONK
instances (gates, copy constraints, grand product) whose quotient passes
the degree check and satisfies the quotient identity at a random point,
while a broken witness or a wrong permutation fails it."""
import random

import pytest

from oracle import plonk as P
from oracle.field import R_MOD, fr_inv
from oracle.ntt import omega_for_size, poly_eval
from tests.plonk_util import shift_for, valid_instance


def test_batch_inv():
    rng = random.Random(1)
    xs = [rng.randrange(R_MOD) for _ in range(50)] + [0, 0, 1, R_MOD - 1]
    rng.shuffle(xs)
    assert P.batch_inv(xs) == [pow(x, -1, R_MOD) if x else 0 for x in xs]
    assert P.batch_inv([]) == [] and P.batch_inv([0, 0]) == [0, 0]


@pytest.mark.parametrize("n", [4, 8])
def test_pipeline_against_horner(n):
    rng = random.Random(n)
    coeffs = {k: [rng.randrange(R_MOD) for _ in range(n)] for k in P.Q_INPUTS if k not in ("zshift", "l1")}
    omega = omega_for_size(n)
    shift = shift_for(n, n)
    kw = dict(n=n, omega=omega, k1=7, k2=13, alpha=rng.randrange(R_MOD), beta=rng.randrange(R_MOD),
              gamma=rng.randrange(R_MOD), shift=shift)
    t_full = P.quotient_full(coeffs, **kw)
    m = 4 * n
    wm = omega_for_size(m)
    for i in range(m):
        x = shift * pow(wm, i, R_MOD) % R_MOD
        ext = {k: [poly_eval(coeffs[k], x)] for k in coeffs}
        ext["zshift"] = [poly_eval(coeffs["z"], x * omega % R_MOD)]
        ext["l1"] = [(pow(x, n, R_MOD) - 1) * fr_inv(n * (x - 1) % R_MOD) % R_MOD]  # L1 in closed form
        num, _ = P.quotient_evals(ext, n=n, m=1, omega_m=wm, shift=x, shift_n=pow(x, n, R_MOD),
                                  alpha=kw["alpha"], beta=kw["beta"], gamma=kw["gamma"], k1=7, k2=13)
        assert poly_eval(t_full, x) == num[0] * fr_inv((pow(x, n, R_MOD) - 1) % R_MOD) % R_MOD


@pytest.mark.parametrize("n,seed", [(8, 1), (16, 2), (32, 3)])
def test_valid_instance_quotient(n, seed):
    inst = valid_instance(n, seed)
    kw = {k: inst[k] for k in ("n", "omega", "k1", "k2", "alpha", "beta", "gamma")}
    t1, t2, t3 = P.quotient(inst["coeffs"], shift=shift_for(n, seed), **kw)
    # the quotient identity at a random point zeta
    co = inst["coeffs"]
    zeta = random.Random(seed + 100).randrange(R_MOD)
    ev = lambda k: poly_eval(co[k], zeta)
    a, b, c, z = ev("a"), ev("b"), ev("c"), ev("z")
    zw = poly_eval(co["z"], zeta * inst["omega"] % R_MOD)
    al, be, ga = inst["alpha"], inst["beta"], inst["gamma"]
    gate = (ev("ql") * a + ev("qr") * b + ev("qm") * a * b + ev("qo") * c + ev("qc") + ev("pi")) % R_MOD
    perm = ((a + be * zeta + ga) * (b + be * 7 * zeta + ga) * (c + be * 13 * zeta + ga) * z
            - (a + be * ev("s1") + ga) * (b + be * ev("s2") + ga) * (c + be * ev("s3") + ga) * zw) % R_MOD
    l1 = (pow(zeta, n, R_MOD) - 1) * fr_inv(n * (zeta - 1) % R_MOD) % R_MOD
    num = (gate + al * perm + al * al % R_MOD * (z - 1) * l1) % R_MOD
    t = (poly_eval(t1, zeta) + pow(zeta, n, R_MOD) * poly_eval(t2, zeta)
         + pow(zeta, 2 * n, R_MOD) * poly_eval(t3, zeta)) % R_MOD
    assert t * (pow(zeta, n, R_MOD) - 1) % R_MOD == num


def test_broken_instances_fail_degree_check():
    n = 16
    inst = valid_instance(n, 9)
    kw = {k: inst[k] for k in ("n", "omega", "k1", "k2", "alpha", "beta", "gamma")}
    from oracle.ntt import coeffs_from_evals_on_roots
    for key, row in (("c", 3), ("s1", 5), ("z", 7)):   # a gate, the permutation, the grand product
        ev = list(inst["evals"][key])
        ev[row] = (ev[row] + 1) % R_MOD
        co = dict(inst["coeffs"])
        co[key] = coeffs_from_evals_on_roots(ev, inst["omega"])
        with pytest.raises(ValueError, match="quotient degree too large"):
            P.quotient(co, shift=shift_for(n, 9), **kw)
