"""Host G2 (libzkhip CPU code) against the oracle.  CPU only (no device)."""
import json
import os
import random

import pytest

from oracle import curve2 as O
from oracle.field import R_MOD

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pyzkp_amd", "libzkhip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libzkhip.so not built")


def _aff(p):
    from pyzkp_amd.g2 import normalize
    return normalize(p)


def test_oracle_g2_pins():
    assert O.is_on_twist(O.G2_GENERATOR)
    assert O.mul(O.G2_GENERATOR, R_MOD) is None
    with open(os.path.join(os.path.dirname(__file__), "golden", "plonk_pk_n16.json")) as f:
        pk = json.load(f)
    g0, g1 = (tuple(tuple(c) for c in p) for p in pk["srs_g2"])
    assert g0 == O.G2_GENERATOR
    assert O.is_on_twist(g1)
    a, b = 12345, 67890
    assert O.add(O.mul(g1, a), O.mul(g1, b)) == O.mul(g1, a + b)


def test_g2_mul_add_match_oracle():
    from pyzkp_amd import g2 as G
    rng = random.Random(1)
    G0 = (O.G2_GENERATOR[0], O.G2_GENERATOR[1], (1, 0))
    for _ in range(5):
        k1, k2 = rng.randrange(R_MOD), rng.randrange(R_MOD)
        p1, p2 = G.g2_mul(G0, k1), G.g2_mul(G0, k2)
        assert _aff(p1) == O.mul(O.G2_GENERATOR, k1)
        assert _aff(G.g2_add(p1, p2)) == O.add(O.mul(O.G2_GENERATOR, k1), O.mul(O.G2_GENERATOR, k2))
        assert _aff(G.g2_add(p1, p1)) == O.mul(O.G2_GENERATOR, 2 * k1)
        neg = (p1[0], ((-p1[1][0]) % O.Q, (-p1[1][1]) % O.Q), (1, 0))
        assert _aff(G.g2_add(p1, neg)) is None
    assert _aff(G.g2_mul(G0, 0)) is None
    assert _aff(G.g2_mul(G0, R_MOD)) is None


@pytest.mark.parametrize("n,c", [(1, 0), (7, 3), (64, 0), (200, 6)])
def test_g2_msm_matches_oracle(n, c):
    from pyzkp_amd import g2 as G
    rng = random.Random(n)
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    ss = [rng.randrange(R_MOD) for _ in range(n)]
    G0 = (O.G2_GENERATOR[0], O.G2_GENERATOR[1], (1, 0))
    words = G.g2_mul_many(G0, ks)
    bases = G.from_words(words)
    got = _aff(G.g2_msm(bases, ss, window_bits=c, threads=4))
    expect = O.mul(O.G2_GENERATOR, sum(k * s for k, s in zip(ks, ss)) % R_MOD)
    assert got == expect
    if n <= 7:
        pts = [O.mul(O.G2_GENERATOR, k) for k in ks]
        assert O.msm_naive(pts, ss) == expect


def test_g2_msm_edge_cases():
    from pyzkp_amd import g2 as G
    G0 = (O.G2_GENERATOR[0], O.G2_GENERATOR[1], (1, 0))
    P = G.g2_mul(G0, 77)
    pts = [P, P, P, G.G2_ZERO, G0]
    ss = [1, 2, R_MOD - 1, 9, 0]
    assert _aff(G.g2_msm(pts, ss, window_bits=4)) == O.mul(O.G2_GENERATOR, 77 * 2 % R_MOD)
    assert _aff(G.g2_msm(pts, [0] * 5)) is None


@pytest.mark.parametrize("c", [3, 5, 8])
def test_g2_msm_batches_and_collisions(c):
    # more points than one batch of affine additions, few buckets: many
    # additions meet a bucket that is busy in the current batch; repeated
    # and negated bases put P + P and P + (-P) into one bucket
    from pyzkp_amd import g2 as G
    rng = random.Random(c)
    n = 2500
    base_ks = [rng.randrange(1, R_MOD) for _ in range(40)]
    ks = [base_ks[rng.randrange(40)] * (1 if rng.random() < 0.7 else -1) % R_MOD for _ in range(n)]
    ss = [rng.randrange(R_MOD) if rng.random() < 0.8 else rng.randrange(1 << 10) for _ in range(n)]
    G0 = (O.G2_GENERATOR[0], O.G2_GENERATOR[1], (1, 0))
    bases = G.from_words(G.g2_mul_many(G0, ks))
    expect = O.mul(O.G2_GENERATOR, sum(k * s for k, s in zip(ks, ss)) % R_MOD)
    for threads in (1, 3, 8):
        assert _aff(G.g2_msm(bases, ss, window_bits=c, threads=threads)) == expect
    # every scalar equal: all points of a window land in one bucket
    s = rng.randrange(R_MOD)
    expect2 = O.mul(O.G2_GENERATOR, sum(ks) * s % R_MOD)
    assert _aff(G.g2_msm(bases, [s] * n, window_bits=c, threads=4)) == expect2


def test_g2_chord_extreme_coordinates():
    # The Fq2 products reduce lazily (512-bit Karatsuba terms, q^2 added to
    # keep c0 non-negative); coefficients at 0, 1 and next to q are their
    # extreme inputs.  The chord rule (x3, y3) = (l^2 - x1 - x2,
    # l (x1 - x3) - y1), l = (y2 - y1) / (x2 - x1), does not use the curve
    # equation, so off-curve points with such coordinates check every
    # product against the oracle's affine formula: through the Jacobian
    # mixed addition (g2_add) and through the MSM's batched affine bucket
    # additions (two unit scalars).
    from pyzkp_amd import g2 as G
    from oracle.field import Q_MOD
    rng = random.Random(0x6262)
    edge = [0, 1, 2, Q_MOD - 1, Q_MOD - 2, (Q_MOD - 1) // 2, (Q_MOD + 1) // 2, (1 << 253) - 1]
    pick = lambda: rng.choice(edge) if rng.random() < 0.75 else rng.randrange(Q_MOD)
    checked = 0
    while checked < 60:
        p = ((pick(), pick()), (pick(), pick()))
        q = ((pick(), pick()), (pick(), pick()))
        if p[0] == q[0] or p == ((0, 0), (0, 0)) or q == ((0, 0), (0, 0)):
            continue
        expect = O.add(p, q)
        P, Q = (p[0], p[1], (1, 0)), (q[0], q[1], (1, 0))
        assert _aff(G.g2_add(P, Q)) == expect, (p, q)
        assert _aff(G.g2_msm([P, Q], [1, 1], window_bits=4, threads=1)) == expect, (p, q)
        checked += 1


def _simd_modes():
    from pyzkp_amd._native import load_library
    lib = load_library()
    modes = [0]
    prev = lib.zk_host_simd(-1)
    if lib.zk_host_simd(1):
        modes.append(1)
    lib.zk_host_simd(prev)
    return lib, modes


@pytest.mark.parametrize("c", [4, 9, 12])
def test_g2_msm_ifma_and_scalar_paths_agree(c):
    # The AVX-512 IFMA accumulation (where the CPU has it) and the 64-bit
    # scalar one, on uniform scalars and on skewed ones (zeros, small values,
    # repeated and negated bases, a block of equal scalars that piles one
    # bucket per window), both against the closed form sum k_i s_i G.
    from pyzkp_amd import g2 as G
    lib, modes = _simd_modes()
    rng = random.Random(0x1F3A + c)
    n = 3000
    base_ks = [rng.randrange(1, R_MOD) for _ in range(n // 3)]
    ks = [base_ks[i % len(base_ks)] * (1 if rng.random() < 0.8 else -1) % R_MOD for i in range(n)]
    G0 = (O.G2_GENERATOR[0], O.G2_GENERATOR[1], (1, 0))
    bases = G.from_words(G.g2_mul_many(G0, ks))
    uniform = [rng.randrange(R_MOD) for _ in range(n)]
    skewed = []
    for i in range(n):
        u = rng.random()
        skewed.append(0 if u < 0.3 else rng.randrange(1 << 16) if u < 0.5 else 77 if i % 7 == 0
                      else rng.randrange(R_MOD))
    try:
        for ss in (uniform, skewed):
            expect = O.mul(O.G2_GENERATOR, sum(k * s for k, s in zip(ks, ss)) % R_MOD)
            for mode in modes:
                lib.zk_host_simd(mode)
                assert _aff(G.g2_msm(bases, ss, window_bits=c, threads=3)) == expect, (mode, c)
    finally:
        lib.zk_host_simd(1)


def test_host_simd_switch():
    lib, modes = _simd_modes()
    assert lib.zk_host_simd(0) == 0
    assert lib.zk_host_simd(-1) == 0
    assert lib.zk_host_simd(1) == (1 if 1 in modes else 0)
    assert lib.zk_host_simd(7) == lib.zk_host_simd(-1)


def test_g2_prepared_sets():
    # A prepared set (zk_g2_bases_create) gives the same points as the
    # per-call conversion, on either arithmetic path, for prefixes of the
    # set; n beyond the set and null handles are refused; an empty set
    # gives the identity.
    import ctypes as ct
    import numpy as np
    from pyzkp_amd import g2 as G
    lib, modes = _simd_modes()
    rng = random.Random(0x77A)
    n = 700
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    ss = [rng.randrange(R_MOD) for _ in range(n)]
    G0 = (O.G2_GENERATOR[0], O.G2_GENERATOR[1], (1, 0))
    words = G.g2_mul_many(G0, ks)
    words[5] = 0                                    # a point at infinity inside the set
    bases = G.from_words(words)
    sw = np.frombuffer(b"".join(s.to_bytes(32, "little") for s in ss), dtype="<u8").reshape(n, 4).copy()
    try:
        for mode in modes:
            lib.zk_host_simd(mode)
            for m in (1, 333, n):
                a = G.g2_msm_words(bases, sw, m, window_bits=7, threads=3)
                b = G.g2_msm_words(bases, sw, m, window_bits=7, threads=3, prepared=False)
                assert _aff(a) == _aff(b), (mode, m)
            expect = O.mul(O.G2_GENERATOR, sum(k * s for i, (k, s) in enumerate(zip(ks, ss)) if i != 5) % R_MOD)
            assert _aff(G.g2_msm_words(bases, sw, n, threads=4)) == expect
    finally:
        lib.zk_host_simd(1)
    out = np.zeros(16, dtype="<u8")
    h = bases.prepared()
    assert lib.zk_g2_msm_prepared(h, sw.ctypes.data, n + 1, 0, 2, out.ctypes.data) == 1
    assert lib.zk_g2_msm_prepared(None, sw.ctypes.data, 1, 0, 2, out.ctypes.data) == 1
    assert lib.zk_g2_msm_prepared(h, sw.ctypes.data, 0, 0, 2, out.ctypes.data) == 0 and not out.any()
    e = ct.c_void_p()
    assert lib.zk_g2_bases_create(None, 0, 1, ct.byref(e)) == 0 and e.value
    assert lib.zk_g2_msm_prepared(e, sw.ctypes.data, 0, 0, 1, out.ctypes.data) == 0 and not out.any()
    assert lib.zk_g2_bases_destroy(e) == 0
    assert lib.zk_g2_bases_destroy(None) == 0
    bases.free()
    bases.free()                                    # idempotent; the next MSM prepares again
    assert _aff(G.g2_msm_words(bases, sw, n, threads=2)) == expect


@pytest.mark.parametrize("c", [3, 15])
def test_g2_msm_2_15_both_paths(c):
    # 2^15 points: with c = 3 (4 buckets a window) every window's additions
    # pile into few buckets, so the IFMA path's replica pool fills and is
    # folded back many times; c = 15 has more buckets than points per window
    # and most buckets hold one point.  Both paths against the closed form.
    import numpy as np
    from pyzkp_amd import g2 as G
    lib, modes = _simd_modes()
    rng = random.Random(0x215 + c)
    n = 1 << 15
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    ss = [rng.randrange(R_MOD) for _ in range(n)]
    G0 = (O.G2_GENERATOR[0], O.G2_GENERATOR[1], (1, 0))
    bases = G.from_words(G.g2_mul_many(G0, ks))
    sw = np.frombuffer(b"".join(s.to_bytes(32, "little") for s in ss), dtype="<u8").reshape(n, 4).copy()
    expect = O.mul(O.G2_GENERATOR, sum(k * s for k, s in zip(ks, ss)) % R_MOD)
    try:
        for mode in modes:
            lib.zk_host_simd(mode)
            assert _aff(G.g2_msm_words(bases, sw, n, window_bits=c, threads=8)) == expect, mode
    finally:
        lib.zk_host_simd(1)


def test_g2_bucket_sum_exceptional_cases():
    # With c = 6 a window has 32 buckets, eight per segment of the IFMA
    # path's lane-parallel bucket sum, so small digits place points in
    # neighbouring buckets of one segment and drive each exceptional case of
    # its running sums: run + B with B = run (doubling) and B = -run
    # (cancellation), sum + run with sum = run (doubling) and sum = -run
    # (cancellation).  Both paths against the closed form.
    from pyzkp_amd import g2 as G
    lib, modes = _simd_modes()
    G0 = (O.G2_GENERATOR[0], O.G2_GENERATOR[1], (1, 0))
    k = 0x1234567
    P = G.g2_mul(G0, k)
    neg = lambda p: (p[0], ((-p[1][0]) % O.Q, (-p[1][1]) % O.Q), (1, 0))
    cases = [
        ([P, P], [1, 2], [k, k]),                          # bucket 1 = P, bucket 0 = P: run doubles
        ([P, neg(P)], [2, 1], [k, -k]),                    # bucket 0 = -P cancels run = P
        ([P], [2], [k]),                                   # empty bucket 0: sum = run doubles
        ([P, G.g2_mul(G0, (-2 * k) % R_MOD)], [3, 2], [k, -2 * k]),  # run = -P meets sum = P
        ([P, P, neg(P), G.g2_mul(G0, 3 * k)], [5, 4, 3, 1], [k, k, -k, 3 * k]),
    ]
    try:
        for pts, ss, ks in cases:
            expect = O.mul(O.G2_GENERATOR, sum(a * b for a, b in zip(ks, ss)) % R_MOD)
            for mode in modes:
                lib.zk_host_simd(mode)
                for threads in (1, 4):
                    assert _aff(G.g2_msm(pts, ss, window_bits=6, threads=threads)) == expect, (ss, mode)
    finally:
        lib.zk_host_simd(1)
