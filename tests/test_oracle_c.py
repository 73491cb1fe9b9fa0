"""Pins the C oracle against the Python oracle and closed forms.  CPU only."""
import random

import pytest

from oracle import cref
from oracle import curve as C
from oracle import msm as M
from oracle import ntt as N
from oracle.field import R_MOD


@pytest.mark.parametrize("log_n", [0, 1, 2, 5, 10])
def test_c_ntt_matches_python(log_n):
    n = 1 << log_n
    rng = random.Random(log_n + 50)
    a = [rng.randrange(R_MOD) for _ in range(n)]
    w = N.omega_for_size(n)
    assert cref.ntt(a, n, w, False) == N.evals_from_coeffs_on_roots(a, n, w)
    assert cref.ntt(a, n, w, True) == N.coeffs_from_evals_on_roots(a, w)
    for s in (5, 0):
        assert cref.ntt(a, n, w, False, shift=s) == N.evals_from_coeffs_on_coset(a, n=n, omega=w, shift=s)
        assert cref.ntt(a, n, w, True, shift=s) == N.coeffs_from_evals_on_coset(a, omega=w, shift=s)
    # pad / truncate
    b = a + [rng.randrange(R_MOD) for _ in range(7)]
    assert cref.ntt(b, n, w, False) == N.evals_from_coeffs_on_roots(b, n, w)
    assert cref.ntt(a[: n // 2], n, w, False) == N.evals_from_coeffs_on_roots(a[: n // 2], n, w)


def test_c_ntt_threads_agree():
    n = 1 << 12
    rng = random.Random(1)
    a = [rng.randrange(R_MOD) for _ in range(n)]
    w = N.omega_for_size(n)
    assert cref.ntt(a, n, w, False, threads=1) == cref.ntt(a, n, w, False, threads=4)


def test_c_ntt_golden(tmp_path):
    import json, os
    with open(os.path.join(os.path.dirname(__file__), "golden", "plonk_pk_n16.json")) as f:
        pk = json.load(f)
    for name, ev in pk["evals"].items():
        assert cref.ntt(ev, 16, pk["omega"], True) == pk["coeffs"][name]


@pytest.mark.parametrize("n,w", [(1, 4), (5, 3), (64, 8), (200, 6), (300, 11)])
def test_c_msm_matches_closed_form(n, w):
    rng = random.Random(n * 7 + w)
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    ss = [rng.randrange(R_MOD) for _ in range(n)]
    pts = [cref.g1_mul(C.G1_GENERATOR, k) for k in ks]
    assert pts[0] == C.g1_mul_affine(C.G1_GENERATOR, ks[0])
    expect = M.msm_closed_form(ks, ss)
    assert cref.msm_pippenger(pts, ss, w) == expect
    assert cref.msm_pippenger(pts, ss, w, threads=4) == expect
    if n <= 64:
        assert M.msm_pippenger(pts, ss, w) == expect


def test_c_msm_edge_cases():
    G = C.G1_GENERATOR
    P = C.g1_mul_affine(G, 77)
    pts = [P, P, P, C.aff_neg(P), G, None]
    ss = [1, 2, R_MOD - 1, 5, 0, 9]
    assert cref.msm_pippenger(pts, ss, 4) == C.g1_mul_affine(G, (77 * (1 + 2 + R_MOD - 1 - 5)) % R_MOD)
    assert cref.msm_pippenger(pts, [0] * 6, 4) is None


def test_c_poly_eval_matches_ntt_rows():
    n = 64
    rng = random.Random(2)
    a = [rng.randrange(R_MOD) for _ in range(n)]
    w = N.omega_for_size(n)
    ev = N.evals_from_coeffs_on_roots(a, n, w)
    import numpy as np
    words = cref._words(a)
    for i in (0, 1, 17, 63):
        assert cref.poly_eval_words(words, n, pow(w, i, R_MOD)) == ev[i]
