"""The device G2 MSM (zk_msm_g2) through the C-ABI, against the closed form
[sum s_i k_i] Q for bases [k_i] Q (Q the G2 generator) and against the host
G2 MSM (zk_g2_msm_host), on sizes from one point up, every window class,
zero scalars, r - 1, identity bases, repeated bases (the doubling case) and
cancelling pairs."""
import random

import numpy as np
import pytest

from oracle.field import R_MOD
from pyzkp_amd import g2 as G2
from pyzkp_amd import workload

pytestmark = pytest.mark.gpu


def _words(xs):
    return np.array([[(x >> (64 * i)) & ((1 << 64) - 1) for i in range(4)] for x in xs], dtype=np.uint64)


def _bases(ks):
    """[k_i] Q as (n, 16) host words (k = 0 gives the identity, all zeros)."""
    return G2.g2_mul_many(G2.G2_GENERATOR, [k % R_MOD for k in ks])


def _closed_form(ks, ss):
    t = sum(k * s for k, s in zip(ks, ss)) % R_MOD
    return G2.g2_mul_many(G2.G2_GENERATOR, [t])[0]


def _device_msm(zctx, words, ss, c):
    gb = zctx.msm_g2_bases_create(words)
    sc = workload.device_fr_mont(zctx, _words(ss))
    try:
        return zctx.msm_g2(gb, sc.ptr, len(ss), c)
    finally:
        sc.free()
        gb.free()


@pytest.mark.parametrize("n", [1, 2, 5, 100, 3000])
def test_msm_g2_closed_form(zctx, n):
    rng = random.Random(0x92 + n)
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    ss = [rng.randrange(R_MOD) for _ in range(n)]
    words = _bases(ks)
    expect = _closed_form(ks, ss)
    for c in (0, 2, 3, 8, 13):
        print(f"G2 MSM n={n} c={c}", flush=True)
        got = _device_msm(zctx, words, ss, c)
        assert np.array_equal(got, expect), (n, c)


def test_msm_g2_edge_cases(zctx):
    # zero scalars, r - 1, identity bases, a base repeated (doubling in a
    # bucket) and a base with its negation under equal scalars (cancelling)
    rng = random.Random(0x93)
    n = 600
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    for i in range(0, n, 7):
        ks[i] = ks[0]                       # repeats
    for i in range(3, n, 11):
        ks[i] = R_MOD - ks[1]               # negations of base 1
    for i in range(5, n, 13):
        ks[i] = 0                           # identity bases
    ss = [rng.randrange(R_MOD) for _ in range(n)]
    for i in range(0, n, 5):
        ss[i] = 0
    for i in range(2, n, 17):
        ss[i] = R_MOD - 1
    for i in range(3, n, 11):
        ss[i] = ss[1]                       # base 1 and its negation, same scalar
    words = _bases(ks)
    expect = _closed_form(ks, ss)
    for c in (0, 4, 10):
        assert np.array_equal(_device_msm(zctx, words, ss, c), expect), c
    # all-zero scalars: the identity
    assert not _device_msm(zctx, words, [0] * n, 0).any()


def test_msm_g2_matches_host_2_16(zctx):
    n = 1 << 16
    kw = workload.fr_uniform(0x94, n)
    sw = workload.fr_uniform(0x95, n)
    from pyzkp_amd.field import words_to_ints
    ks, ss = words_to_ints(kw), words_to_ints(sw)
    words = G2.g2_mul_many_words(G2.G2_GENERATOR, kw)
    host = G2.g2_msm_words(G2.from_words(words), sw, n)
    got = _device_msm(zctx, words, ss, 0)
    assert G2.point_from_words(got) == host
    assert np.array_equal(got, _closed_form(ks, ss))


def test_msm_g2_prefix_empty_and_refused(zctx):
    # a call over the first n bases of a larger set, n = 0 (the identity) and
    # more scalars than bases (refused with ValueError before any device work)
    rng = random.Random(0x96)
    m = 400
    ks = [rng.randrange(1, R_MOD) for _ in range(m)]
    ss = [rng.randrange(R_MOD) for _ in range(m)]
    gb = zctx.msm_g2_bases_create(_bases(ks))
    sc = workload.device_fr_mont(zctx, _words(ss))
    try:
        for n in (1, 77, m):
            print(f"G2 MSM prefix n={n} of {m}", flush=True)
            assert np.array_equal(zctx.msm_g2(gb, sc.ptr, n, 0), _closed_form(ks[:n], ss[:n])), n
        assert not zctx.msm_g2(gb, sc.ptr, 0, 0).any()
        with pytest.raises(ValueError):
            zctx.msm_g2(gb, sc.ptr, m + 1, 0)
        # the context stays usable after the refused call
        assert np.array_equal(zctx.msm_g2(gb, sc.ptr, m, 9), _closed_form(ks, ss))
    finally:
        sc.free()
        gb.free()
