"""PLONK quotient on the device: BATCH_INV and the fused
PLONK_T_QUOTIENT_EVALS operator through the runtime, and the whole quotient
graph (15 coset NTTs at 4n, numerator, inversion, coset iNTT) equal to the
oracle bit for bit on valid instances and on random inputs; a broken
witness fails the degree check; at n = 2^16 the quotient identity holds at
a random point."""
import random

import pytest

from oracle import cref
from oracle import plonk as P
from oracle.field import R_MOD, fr_inv
from oracle.ntt import poly_eval
from pyzkp_amd.plonk import compute_quotient
from pyzkp_amd.runtime import Backend, Device, DType, Executor, Graph, KernelRegistry, OpType, RuntimeConfig
from pyzkp_amd.runtime import register_hip_kernels
from tests.plonk_util import shift_for, valid_instance

pytestmark = pytest.mark.gpu


def _kw(inst):
    return {k: inst[k] for k in ("n", "omega", "k1", "k2", "alpha", "beta", "gamma")}


def test_batch_inv_op(zctx):
    rng = random.Random(3)
    for n in (1, 7, 100, 70000):
        xs = [rng.randrange(R_MOD) if rng.random() < 0.9 else 0 for _ in range(n)]
        g = Graph()
        g.add_buffer(id="x", device=Device.CPU, dtype=DType.FR, data=xs)
        g.add_node(op=OpType.BATCH_INV, inputs=["x"], outputs=["y_d"])
        g.add_node(op=OpType.FROM_DEVICE, inputs=["y_d"], outputs=["y"])
        r = KernelRegistry()
        register_hip_kernels(r)
        Executor(registry=r).run(g, runtime_config=RuntimeConfig(backend=Backend.HIP), keep=["y"])
        assert g.buffers["y"].data == P.batch_inv(xs), n


@pytest.mark.parametrize("n,seed", [(8, 1), (64, 2), (1024, 3)])
def test_quotient_matches_oracle(zctx, n, seed):
    inst = valid_instance(n, seed)
    shift = shift_for(n, seed)
    got = compute_quotient(inst["coeffs"], shift=shift, **_kw(inst))
    assert got == tuple(tuple(t) for t in P.quotient(inst["coeffs"], shift=shift, **_kw(inst)))


def test_quotient_random_inputs_full(zctx):
    n = 32
    rng = random.Random(7)
    coeffs = {k: [rng.randrange(R_MOD) for _ in range(n)] for k in P.Q_INPUTS if k not in ("zshift", "l1")}
    from oracle.ntt import omega_for_size
    kw = dict(n=n, omega=omega_for_size(n), k1=7, k2=13, alpha=rng.randrange(R_MOD), beta=rng.randrange(R_MOD),
              gamma=rng.randrange(R_MOD))
    shift = shift_for(n, 7)
    assert compute_quotient(coeffs, shift=shift, check_degree=False, **kw) == P.quotient_full(coeffs, shift=shift,
                                                                                              **kw)
    with pytest.raises(ValueError, match="quotient degree too large"):
        compute_quotient(coeffs, shift=shift, **kw)


def test_quotient_large_identity(zctx):
    n = 1 << 16
    inst = valid_instance(n, 11, intt=lambda v, w: cref.ntt(v, len(v), w, True, threads=0))
    t1, t2, t3 = compute_quotient(inst["coeffs"], shift=shift_for(n, 11), **_kw(inst))
    co = inst["coeffs"]
    zeta = random.Random(12).randrange(R_MOD)
    ev = lambda k: poly_eval(co[k], zeta)
    a, b, c, z = ev("a"), ev("b"), ev("c"), ev("z")
    zw = poly_eval(co["z"], zeta * inst["omega"] % R_MOD)
    al, be, ga = inst["alpha"], inst["beta"], inst["gamma"]
    gate = (ev("ql") * a + ev("qr") * b + ev("qm") * a * b + ev("qo") * c + ev("qc") + ev("pi")) % R_MOD
    perm = ((a + be * zeta + ga) * (b + be * 7 * zeta + ga) * (c + be * 13 * zeta + ga) * z
            - (a + be * ev("s1") + ga) * (b + be * ev("s2") + ga) * (c + be * ev("s3") + ga) * zw) % R_MOD
    l1 = (pow(zeta, n, R_MOD) - 1) * fr_inv(n * (zeta - 1) % R_MOD) % R_MOD
    num = (gate + al * perm + al * al % R_MOD * (z - 1) * l1) % R_MOD
    zn = pow(zeta, n, R_MOD)
    t = (poly_eval(t1, zeta) + zn * poly_eval(t2, zeta) + zn * zn % R_MOD * poly_eval(t3, zeta)) % R_MOD
    assert t * (zn - 1) % R_MOD == num
