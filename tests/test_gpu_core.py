"""GPU parity of the NTT and MSM kernels against the oracle, through the
C-ABI (ctypes Context)."""
import random

import pytest

from oracle import curve as C
from oracle import msm as M
from oracle import ntt as N
from oracle.field import R_MOD, Q_MOD

from tests.gpu_util import download_fr, msm_affine, upload_bases, upload_fr_mont

pytestmark = pytest.mark.gpu


def _mont(x):
    return x * (1 << 256) % R_MOD


@pytest.mark.parametrize("log_n", list(range(0, 13)) + [14, 16])
def test_ntt_forward_inverse(zctx, log_n):
    n = 1 << log_n
    rng = random.Random(log_n)
    a = [rng.randrange(R_MOD) for _ in range(n)]
    w = N.omega_for_size(n)
    src = upload_fr_mont(zctx, a)
    dst = zctx.alloc(32 * n)
    zctx.ntt(src.ptr, n, dst.ptr, log_n, False)
    assert download_fr(zctx, dst.ptr, n) == N.evals_from_coeffs_on_roots(a, n, w)
    zctx.ntt(src.ptr, n, dst.ptr, log_n, True)
    assert download_fr(zctx, dst.ptr, n) == N.coeffs_from_evals_on_roots(a, w)
    for s in (5, 0):
        zctx.ntt(src.ptr, n, dst.ptr, log_n, False, shift_mont=_mont(s))
        assert download_fr(zctx, dst.ptr, n) == N.evals_from_coeffs_on_coset(a, n=n, omega=w, shift=s)
        zctx.ntt(src.ptr, n, dst.ptr, log_n, True, shift_mont=_mont(s))
        assert download_fr(zctx, dst.ptr, n) == N.coeffs_from_evals_on_coset(a, omega=w, shift=s)
    src.free()
    dst.free()


@pytest.mark.parametrize("log_n,in_len", [(4, 5), (4, 40), (10, 1000), (10, 3000), (13, 1), (13, 0)])
def test_ntt_pad_truncate(zctx, log_n, in_len):
    n = 1 << log_n
    rng = random.Random(in_len)
    a = [rng.randrange(R_MOD) for _ in range(in_len)]
    src = upload_fr_mont(zctx, a)
    dst = zctx.alloc(32 * n)
    zctx.ntt(src.ptr, in_len, dst.ptr, log_n, False)
    assert download_fr(zctx, dst.ptr, n) == N.evals_from_coeffs_on_roots(a, n, N.omega_for_size(n))


@pytest.mark.parametrize("log_n", [3, 11, 15])
def test_ntt_in_place_and_pass_splits(zctx, log_n):
    n = 1 << log_n
    rng = random.Random(99)
    a = [rng.randrange(R_MOD) for _ in range(n)]
    w = N.omega_for_size(n)
    expect = N.evals_from_coeffs_on_roots(a, n, w)
    for bits in (0, 1, 2, 3, 5, 7, 11):
        zctx.set_tuning("ntt_max_pass_bits", bits)
        buf = upload_fr_mont(zctx, a)
        zctx.ntt(buf.ptr, n, buf.ptr, log_n, False)
        assert download_fr(zctx, buf.ptr, n) == expect, bits
        buf.free()
    zctx.set_tuning("ntt_max_pass_bits", 0)


@pytest.mark.parametrize("inplace,stage,full", [(1, 1, 1), (0, 1, 1), (1, 0, 1), (0, 0, 0), (1, 1, 0)])
def test_ntt_pass_forms(zctx, inplace, stage, full):
    # in-place or Stockham-order pass stages, per-stage or single twiddle
    # tables, coset scaling from full or two-level tables: all four
    # variants against the oracle, one- and multi-pass plans, odd pass sizes
    # (a lone radix-2 stage), many-column tiles
    import numpy as np
    from oracle import cref
    from pyzkp_amd.field import words_to_ints
    zctx.set_tuning("ntt_inplace", inplace)
    zctx.set_tuning("ntt_stage_twiddles", stage)
    zctx.set_tuning("ntt_full_scale", full)
    try:
        for log_n, bits in ((5, 0), (11, 0), (13, 5), (17, 0), (17, 9)):
            n = 1 << log_n
            rng = random.Random(log_n * 7 + bits)
            a = [rng.randrange(R_MOD) for _ in range(n)]
            aw = np.array([[(x >> (64 * i)) & ((1 << 64) - 1) for i in range(4)] for x in a], dtype=np.uint64)
            w = N.omega_for_size(n)
            src = upload_fr_mont(zctx, a)
            dst = zctx.alloc(32 * n)
            zctx.set_tuning("ntt_max_pass_bits", bits)
            for inv, sh in ((False, None), (True, None), (False, 5), (True, 5)):
                zctx.ntt(src.ptr, n, dst.ptr, log_n, inv, shift_mont=_mont(sh) if sh else None)
                expect = words_to_ints(cref.ntt_words(aw.reshape(-1), n, log_n, w, inv, sh, threads=0).reshape(-1, 4))
                assert download_fr(zctx, dst.ptr, n) == expect, (log_n, bits, inv, sh)
            src.free()
            dst.free()
    finally:
        zctx.set_tuning("ntt_max_pass_bits", 0)
        zctx.set_tuning("ntt_inplace", 1)
        zctx.set_tuning("ntt_stage_twiddles", 1)
        zctx.set_tuning("ntt_full_scale", 1)


def test_ntt_batch(zctx):
    log_n, batch = 9, 5
    n = 1 << log_n
    rng = random.Random(3)
    rows = [[rng.randrange(R_MOD) for _ in range(n)] for _ in range(batch)]
    src = upload_fr_mont(zctx, [x for r in rows for x in r])
    dst = zctx.alloc(32 * n * batch)
    zctx.ntt_batch(src.ptr, dst.ptr, log_n, batch, False, shift_mont=_mont(5))
    out = download_fr(zctx, dst.ptr, n * batch)
    w = N.omega_for_size(n)
    for b in range(batch):
        assert out[b * n:(b + 1) * n] == N.evals_from_coeffs_on_coset(rows[b], n=n, omega=w, shift=5)


def test_ntt_errors(zctx):
    buf = zctx.alloc(64)
    with pytest.raises(ValueError):
        zctx.ntt(buf.ptr, 2, buf.ptr, 29, False)
    with pytest.raises(ValueError):
        zctx.ntt(buf.ptr, 1, buf.ptr, 1, True)  # inverse needs in_len == n
    with pytest.raises(ValueError):
        zctx.ntt(buf.ptr, 2, buf.ptr, 1, False, omega_mont=_mont(1))  # not primitive


def _bases(ks):
    return [C.g1_mul_affine(C.G1_GENERATOR, k) for k in ks]


@pytest.mark.parametrize("n", [1, 2, 3, 31, 64, 100, 257])
def test_msm_closed_form(zctx, n):
    rng = random.Random(1000 + n)
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    ss = [rng.randrange(R_MOD) for _ in range(n)]
    bases = upload_bases(zctx, _bases(ks))
    sc = upload_fr_mont(zctx, ss)
    expect = M.msm_closed_form(ks, ss)
    for c in (0, 2, 4, 7, 10, 13, 16):
        assert msm_affine(zctx, bases, sc, n, c) == expect, c


@pytest.mark.parametrize("split", [1, 0])
def test_msm_both_reductions(zctx, split):
    # digit-split and chunked bucket reductions, every window size with one
    # or two index bits (c = 2, 3) up to the widest, on sparse windows (most
    # buckets empty) and on dense ones
    rng = random.Random(77)
    for n in (5, 600):
        ks = [rng.randrange(1, R_MOD) for _ in range(n)]
        ss = [rng.randrange(R_MOD) for _ in range(n)]
        bases = upload_bases(zctx, _bases(ks))
        sc = upload_fr_mont(zctx, ss)
        expect = M.msm_closed_form(ks, ss)
        try:
            zctx.set_tuning("msm_reduce_split", split)
            for c in (2, 3, 4, 5, 6, 9, 12, 17, 21):
                assert msm_affine(zctx, bases, sc, n, c) == expect, (n, c)
        finally:
            zctx.set_tuning("msm_reduce_split", 1)


def test_msm_edge_cases(zctx):
    G = C.G1_GENERATOR
    P = C.g1_mul_affine(G, 77)
    pts = [P, P, P, C.aff_neg(P), G, None, P]
    ss = [1, 2, R_MOD - 1, 5, 0, 9, 0]
    expect = C.g1_mul_affine(G, (77 * (1 + 2 + R_MOD - 1 - 5)) % R_MOD)
    bases = upload_bases(zctx, pts)
    sc = upload_fr_mont(zctx, ss)
    for c in (2, 3, 8, 16):
        assert msm_affine(zctx, bases, sc, len(pts), c) == expect
    # all-zero scalars and cancellation to the identity
    sc0 = upload_fr_mont(zctx, [0] * 7)
    assert msm_affine(zctx, bases, sc0, 7, 5) is None
    sc2 = upload_fr_mont(zctx, [1, 0, 0, 1, 0, 0, 0])
    assert msm_affine(zctx, bases, sc2, 7, 5) is None
    # the same cases on a prepared (converted once) base set
    from pyzkp_amd.curve import from_jacobian_mont
    mb = zctx.msm_bases_create(bases.ptr, len(pts))
    aff = lambda r: None if r[2] == 0 else (r[0], r[1])
    for c in (2, 8, 16):
        assert aff(from_jacobian_mont(*zctx.msm_g1_prepared(mb, sc.ptr, len(pts), c))) == expect
    assert aff(from_jacobian_mont(*zctx.msm_g1_prepared(mb, sc2.ptr, 7, 5))) is None
    assert zctx.msm_g1_prepared(mb, sc.ptr, 0, 0)[2] == 0          # empty MSM: the identity
    mb.free()


def test_msm_same_scalar_everywhere(zctx):
    # one bucket per window holds every point: segments span many threads
    n = 3000
    rng = random.Random(4)
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    s = rng.randrange(R_MOD)
    bases = upload_bases(zctx, _bases(ks))
    sc = upload_fr_mont(zctx, [s] * n)
    for seg in (1, 3, 64, 80):
        zctx.set_tuning("msm_segment", seg)
        assert msm_affine(zctx, bases, sc, n, 10) == M.msm_closed_form(ks, [s] * n)
    zctx.set_tuning("msm_segment", 0)  # the library default: chosen per launch size


def test_msm_window_groups(zctx):
    # windows split into groups pipelined over two streams (tuning knob)
    from pyzkp_amd import workload
    from pyzkp_amd.field import words_to_ints
    n = 1 << 16
    kw = workload.fr_uniform(0x61, n)
    sw = workload.fr_uniform(0x62, n)
    bases = workload.device_bases(zctx, kw)
    sc = workload.device_fr_mont(zctx, sw)
    expect = M.msm_closed_form(words_to_ints(kw), words_to_ints(sw))
    from pyzkp_amd.curve import from_jacobian_mont
    mb = zctx.msm_bases_create(bases.ptr, n)
    aff = lambda r: None if r[2] == 0 else (r[0], r[1])
    try:
        for g in (2, 3, 5):
            zctx.set_tuning("msm_groups", g)
            for c in (8, 13):
                assert msm_affine(zctx, bases, sc, n, c) == expect, (g, c)
            # prepared base set, with the digit-split and the chunked reductions
            for split in (1, 0):
                zctx.set_tuning("msm_reduce_split", split)
                assert aff(from_jacobian_mont(*zctx.msm_g1_prepared(mb, sc.ptr, n, 11))) == expect, (g, split)
            zctx.set_tuning("msm_reduce_split", 1)
    finally:
        zctx.set_tuning("msm_groups", 0)
        zctx.set_tuning("msm_reduce_split", 1)
        mb.free()


def test_msm_sort_fused_and_chunked(zctx):
    # bucket-sort segments placed whole by one workgroup next to segments
    # too large for one (40% of the scalars share one value, so that value's
    # bucket fills its segment in every window), and windows whose fine keys
    # exceed the fused kernel's table (c = 20: 2^11 fine keys); fused on and
    # off, both against the closed form
    from pyzkp_amd import workload
    from pyzkp_amd.field import words_to_ints
    n = 1 << 18
    kw = workload.fr_uniform(0x71, n)
    sw = workload.fr_uniform(0x72, n)
    u = workload.uniform01(0x73, n)
    sw[u < 0.4] = sw[0]
    bases = workload.device_bases(zctx, kw)
    sc = workload.device_fr_mont(zctx, sw)
    expect = M.msm_closed_form(words_to_ints(kw), words_to_ints(sw))
    try:
        for fused in (1, 0):
            zctx.set_tuning("msm_sort_fused", fused)
            for c in (5, 10, 16, 17, 19, 20):
                assert msm_affine(zctx, bases, sc, n, c) == expect, (fused, c)
    finally:
        zctx.set_tuning("msm_sort_fused", 1)


def test_msm_wide_windows(zctx):
    # c >= 20: up to 2^13 fine keys per coarse key in the bucket sort
    n = 100
    rng = random.Random(20)
    ks = [rng.randrange(1, R_MOD) for _ in range(n)]
    ss = [rng.randrange(R_MOD) for _ in range(n)]
    bases = upload_bases(zctx, _bases(ks))
    sc = upload_fr_mont(zctx, ss)
    for c in (20, 22, 24):
        assert msm_affine(zctx, bases, sc, n, c) == M.msm_closed_form(ks, ss), c


def test_msm_one_huge_bucket(zctx):
    # every point in one bucket of every window: the fine level of the
    # bucket sort spreads that segment over several workgroups
    from pyzkp_amd import workload
    from pyzkp_amd.field import words_to_ints
    n = 1 << 17
    kw = workload.fr_uniform(0x51, n)
    bases = workload.device_bases(zctx, kw)
    s = random.Random(5).randrange(R_MOD)
    sc = upload_fr_mont(zctx, [s] * n)
    expect = M.msm_closed_form(words_to_ints(kw), [s] * n)
    for c in (14, 16, 18):
        assert msm_affine(zctx, bases, sc, n, c) == expect, c


def test_msm_golden_commitments(zctx):
    import json, os
    with open(os.path.join(os.path.dirname(__file__), "golden", "plonk_pk_n16.json")) as f:
        pk = json.load(f)
    srs = [tuple(p) for p in pk["srs_g1"]]
    bases = upload_bases(zctx, srs)
    for name, coeffs in pk["coeffs"].items():
        sc = upload_fr_mont(zctx, coeffs)
        expect = tuple(pk["commits"][name])
        assert msm_affine(zctx, bases, sc, len(coeffs)) == expect, name


def test_ntt_table_cap_evicts_and_rebuilds(zctx):
    """With a 1 MiB cap on cached NTT tables every new size evicts the least
    recently used tables; transforms stay exact when their tables are
    rebuilt."""
    zctx.set_tuning("ntt_table_cap_mb", 1)
    try:
        for rep in range(2):
            for log_n in (14, 12, 15, 13):
                n = 1 << log_n
                rng = random.Random(900 + log_n + rep)
                a = [rng.randrange(R_MOD) for _ in range(n)]
                w = N.omega_for_size(n)
                src = upload_fr_mont(zctx, a)
                dst = zctx.alloc(32 * n)
                zctx.ntt(src.ptr, n, dst.ptr, log_n, False, shift_mont=_mont(5))
                assert download_fr(zctx, dst.ptr, n) == N.evals_from_coeffs_on_coset(a, n=n, omega=w, shift=5)
                zctx.ntt(src.ptr, n, dst.ptr, log_n, True)
                assert download_fr(zctx, dst.ptr, n) == N.coeffs_from_evals_on_roots(a, w)
                src.free()
                dst.free()
    finally:
        zctx.set_tuning("ntt_table_cap_mb", 16384)


def test_batch_plan_tables_survive_table_cap(zctx):
    """Tables a batch plan's recorded launches point at are pinned: a tiny
    table cap and other sizes' transforms in between do not free them."""
    nb = 1 << 10
    rng = random.Random(77)
    fbk = [rng.randrange(1, R_MOD) for _ in range(nb)]
    bases = upload_bases(zctx, _bases(fbk))
    fb = zctx.fixed_bases_create(bases.ptr, nb, 8)
    ss = [rng.randrange(R_MOD) for _ in range(nb)]
    sc = upload_fr_mont(zctx, ss)
    n = 1 << 12
    a = [rng.randrange(R_MOD) for _ in range(n)]
    src, dst = upload_fr_mont(zctx, a), zctx.alloc(32 * n)
    zctx.set_tuning("ntt_table_cap_mb", 1)
    try:
        plan = zctx.batch_create(fb, sc.ptr, [0], [nb], [{"in": src.ptr, "out": dst.ptr, "log_n": 12,
                                                            "batch": 1, "inverse": True}], use_graph=True)
        for log_n in (15, 14, 13):                                     # other tables, evicting
            m = 1 << log_n
            t = upload_fr_mont(zctx, [1] * m)
            zctx.ntt(t.ptr, m, t.ptr, log_n, False)
            zctx.sync()
            t.free()
        for _ in range(2):
            res = plan.run()
            assert download_fr(zctx, dst.ptr, n) == N.coeffs_from_evals_on_roots(a, N.omega_for_size(n))
        from pyzkp_amd.curve import from_jacobian_mont
        p = from_jacobian_mont(*res[0])
        assert (None if p[2] == 0 else (p[0], p[1])) == M.msm_closed_form(fbk, ss)
        plan.free()
    finally:
        zctx.set_tuning("ntt_table_cap_mb", 16384)
    for b in (bases, sc, src, dst):
        b.free()
    fb.free()
