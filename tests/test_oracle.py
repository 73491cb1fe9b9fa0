"""Pins the CPU oracle: golden vectors from the reference's cached PLONK key,
closed forms and brute force.  CPU only."""
import json
import os
import random

import pytest

from oracle import curve as C
from oracle import gen, msm as M, ntt as N
from oracle.field import FR_GENERATOR, Q_MOD, R_MOD, ROOT_2_28

GOLD = os.path.join(os.path.dirname(__file__), "golden", "plonk_pk_n16.json")


@pytest.fixture(scope="module")
def pk():
    with open(GOLD) as f:
        return json.load(f)


# ---------------------------------------------------------------- constants
def test_field_constants():
    for p in (R_MOD, Q_MOD):
        for a in (2, 3, 5, 7):
            assert pow(a, p - 1, p) == 1
    assert (R_MOD - 1) % (1 << 28) == 0 and ((R_MOD - 1) >> 28) % 2 == 1  # 2-adicity 28
    assert pow(FR_GENERATOR, (R_MOD - 1) // 2, R_MOD) == R_MOD - 1  # non-residue
    assert ROOT_2_28 == pow(5, (R_MOD - 1) >> 28, R_MOD)
    assert pow(ROOT_2_28, 1 << 28, R_MOD) == 1
    assert pow(ROOT_2_28, 1 << 27, R_MOD) == R_MOD - 1


# ---------------------------------------------------------------- NTT
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64])
def test_ntt_matches_dft(n):
    rng = random.Random(n)
    a = [rng.randrange(R_MOD) for _ in range(n)]
    w = N.omega_for_size(n)
    assert N.evals_from_coeffs_on_roots(a, n, w) == N.dft_naive(a, w)
    wi = pow(w, -1, R_MOD)
    inv_n = pow(n, -1, R_MOD)
    assert N.coeffs_from_evals_on_roots(a, w) == [x * inv_n % R_MOD for x in N.dft_naive(a, wi)]


def test_ntt_identities():
    n = 32
    w = N.omega_for_size(n)
    delta = [1] + [0] * (n - 1)
    assert N.evals_from_coeffs_on_roots(delta, n, w) == [1] * n
    c = 12345
    assert N.evals_from_coeffs_on_roots([c] * n, n, w) == [c * n % R_MOD] + [0] * (n - 1)
    # truncation and zero padding (crypto/poly/ntt.py:103-110)
    a = list(range(1, 41))
    assert N.evals_from_coeffs_on_roots(a, n, w) == N.dft_naive(a[:n], w)
    assert N.evals_from_coeffs_on_roots(a[:5], n, w) == N.dft_naive(a[:5] + [0] * (n - 5), w)


def test_ntt_rejects_bad_sizes():
    with pytest.raises(ValueError):
        N.omega_for_size(12)
    with pytest.raises(ValueError):
        N.omega_for_size(1 << 29)


def test_coset_matches_horner():
    n = 16
    w = N.omega_for_size(n)
    rng = random.Random(7)
    a = [rng.randrange(R_MOD) for _ in range(n)]
    ev = N.evals_from_coeffs_on_coset(a, n=n, omega=w, shift=5)
    assert ev == [N.poly_eval(a, 5 * pow(w, i, R_MOD) % R_MOD) for i in range(n)]
    assert N.coeffs_from_evals_on_coset(ev, omega=w, shift=5) == a
    # shift 0 keeps out[0] and zeroes the rest (crypto/poly/ntt.py:139-145)
    out = N.coeffs_from_evals_on_coset(ev, omega=w, shift=0)
    assert out[1:] == [0] * (n - 1)
    assert out[0] == N.coeffs_from_evals_on_roots(ev, w)[0]


def test_reference_metal_test_inputs():
    # tests/correctness/metal/test_runtime_metal_fr_ntt.py:20-56 inputs, expected from the DFT
    n = 16
    w = N.omega_for_size(n)
    coeff = [i * i % R_MOD for i in range(n)]
    assert N.evals_from_coeffs_on_roots(coeff, n, w) == N.dft_naive(coeff, w)
    ev = [(i + 3) * 7 % R_MOD for i in range(n)]
    back = N.coeffs_from_evals_on_roots(ev, w)
    assert N.dft_naive(back, w) == ev


def test_golden_intt_vectors(pk):
    n = pk["n"]
    w = pk["omega"]
    assert w == N.omega_for_size(n)
    assert pk["roots"] == [pow(w, i, R_MOD) for i in range(n)]
    for name, ev in pk["evals"].items():
        assert N.coeffs_from_evals_on_roots(ev, w) == pk["coeffs"][name], name
        assert N.evals_from_coeffs_on_roots(pk["coeffs"][name], n, w) == ev, name


# ---------------------------------------------------------------- curve / MSM
def test_curve_law_against_affine_textbook():
    rng = random.Random(11)
    G = C.G1_GENERATOR
    assert C.is_on_curve(G)
    assert C.g1_mul_affine(G, R_MOD) is None
    for _ in range(10):
        a, b = rng.randrange(1, R_MOD), rng.randrange(1, R_MOD)
        P, Qp = C.g1_mul_affine(G, a), C.g1_mul_affine(G, b)
        assert C.is_on_curve(P) and C.is_on_curve(Qp)
        assert C.g1_add_affine(P, Qp) == C.aff_add(P, Qp)
        assert C.g1_add_affine(P, P) == C.aff_add(P, P)
        assert C.g1_add_affine(P, C.aff_neg(P)) is None
        assert C.g1_add_affine(P, Qp) == C.g1_mul_affine(G, (a + b) % R_MOD)


def test_golden_srs_and_commitments(pk):
    srs = [tuple(p) for p in pk["srs_g1"]]
    assert all(C.is_on_curve(p) for p in srs)
    assert srs[0] == C.G1_GENERATOR
    for name, coeffs in pk["coeffs"].items():
        expect = tuple(pk["commits"][name]) if pk["commits"][name] else None
        assert M.msm_naive_g1(srs, coeffs) == expect, name
        assert M.msm_pippenger(srs, coeffs, 4) == expect, name
        assert M.msm_signed_buckets(srs, coeffs, 5) == expect, name


@pytest.mark.parametrize("n", [1, 2, 3, 8, 31, 32, 63, 64, 127])
def test_msm_agreement(n):
    rng = random.Random(100 + n)
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    ss = [rng.randrange(R_MOD) for _ in range(n)]
    pts = [C.g1_mul_affine(C.G1_GENERATOR, k) for k in ks]
    ref = M.msm_closed_form(ks, ss)
    assert M.msm_naive_g1(pts, ss) == ref
    for w in (7, 8):
        assert M.msm_pippenger(pts, ss, w) == ref
    assert M.msm_signed_buckets(pts, ss, 6) == ref


def test_msm_sparse_window_branch():
    # 2^c > 1024 buckets: the gap / scalar-multiple branch
    rng = random.Random(5)
    ks = [rng.randrange(1, R_MOD) for _ in range(12)]
    ss = [rng.randrange(R_MOD) for _ in range(12)]
    pts = [C.g1_mul_affine(C.G1_GENERATOR, k) for k in ks]
    assert M.msm_pippenger(pts, ss, 12) == M.msm_closed_form(ks, ss)


def test_reference_metal_msm_input_set():
    # tests/correctness/metal/test_runtime_metal_g1_msm.py:33-38
    n = 100
    ss = [(i * 12345 + 6789) % R_MOD for i in range(n)]
    pts = [C.g1_mul_affine(C.G1_GENERATOR, i + 1) for i in range(n)]
    expect = C.g1_mul_affine(C.G1_GENERATOR, sum((i + 1) * s for i, s in enumerate(ss)) % R_MOD)
    assert M.msm_pippenger(pts, ss, M.default_window_bits(n)) == expect


def test_repeated_bases_and_edge_scalars():
    G = C.G1_GENERATOR
    P = C.g1_mul_affine(G, 77)
    pts = [P, P, P, C.aff_neg(P), G, None]
    ss = [1, 2, R_MOD - 1, 5, 0, 9]
    expect = C.g1_mul_affine(G, (77 * (1 + 2 + R_MOD - 1 - 5)) % R_MOD)
    assert M.msm_naive_g1(pts, ss) == expect
    assert M.msm_signed_buckets(pts, ss, 3) == expect


@pytest.mark.parametrize("c", [2, 3, 4, 5, 10, 13, 16, 18])
def test_signed_digit_invariants(c):
    W = M.window_count(c)
    rng = random.Random(c)
    ss = [0, 1, R_MOD - 1, (1 << 253), (1 << c) - 1, 1 << (c - 1)] + [rng.randrange(R_MOD) for _ in range(50)]
    for s, row in zip(ss, M.encode_signed_digits(ss, c, W)):
        assert sum(d << (c * k) for k, d in enumerate(row)) == s
        assert all(-(1 << (c - 1)) <= d < (1 << (c - 1)) for d in row)


def test_signed_digits_brute_force_small_c():
    # c = 2, 3: every scalar below 2^12 decomposes exactly
    for c in (2, 3):
        W = 12 // c + 2
        for s in range(1 << 12):
            row = M.encode_signed_digits([s], c, W)[0]
            assert sum(d << (c * k) for k, d in enumerate(row)) == s


def test_csr_against_stable_sort():
    c = 4
    W = M.window_count(c)
    rng = random.Random(9)
    ss = [rng.randrange(R_MOD) for _ in range(40)]
    digits = M.encode_signed_digits(ss, c, W)
    row_ptr, entries = M.csr_buckets(digits, W, c)
    B = (1 << (c - 1)) + 1
    pairs = sorted(((w * B + abs(d), i, d < 0) for i, row in enumerate(digits) for w, d in enumerate(row) if d))
    assert [i | ((1 << 31) if neg else 0) for _, i, neg in pairs] == entries
    assert row_ptr[-1] == len(pairs)


def test_generator_matches_product_copy():
    import numpy as np
    from pyzkp_amd import workload
    from pyzkp_amd.field import words_to_ints

    for seed, n in ((0x1001, 50), (0x2001, 333)):
        assert words_to_ints(workload.fr_uniform(seed, n)) == gen.fr_vector(seed, n)
    rng = gen.SplitMix64(0x3003)
    u = [rng.uniform() for _ in range(20)]
    assert np.allclose(workload.uniform01(0x3003, 20), u, rtol=0, atol=0)
