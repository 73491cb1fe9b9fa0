"""GPU parity at the benchmark sizes: NTT 2^20 bit-exact against the C
oracle, 2^24 by round trip and Horner rows, 2^28 (the largest Fr domain) by a
sparse input's evaluations and a dense round trip; MSM 2^12 (c = 10) against the C
Pippenger and the closed form; 2^16 / 2^22 uniform and skewed by the closed
form [sum s_i k_i] G1 (bases B_i = [k_i] G1)."""
import numpy as np
import pytest

from oracle import cref
from oracle import curve as C
from oracle import msm as M
from oracle import ntt as N
from oracle.field import R_MOD
from pyzkp_amd import workload
from pyzkp_amd.curve import from_jacobian_mont
from pyzkp_amd.field import FR_GENERATOR, fr_to_mont, words_to_ints

pytestmark = pytest.mark.gpu


def _dev_words(ctx, ptr, n):
    """Device Montgomery Fr -> canonical (n, 4) words (converted on device)."""
    tmp = ctx.alloc(32 * n)
    ctx.fr_from_mont(ptr, tmp.ptr, n)
    w = np.frombuffer(ctx.download(tmp.ptr, 32 * n), dtype="<u8").reshape(n, 4).copy()
    tmp.free()
    return w


@pytest.mark.parametrize("variant", ["fwd", "inv", "coset_fwd", "coset_inv"])
def test_ntt_2_20_bit_exact(zctx, variant):
    log_n = 20
    n = 1 << log_n
    a = workload.fr_uniform(0x2001, n)
    src = workload.device_fr_mont(zctx, a)
    dst = zctx.alloc(32 * n)
    inv = variant.endswith("inv")
    shift = FR_GENERATOR if variant.startswith("coset") else None
    zctx.ntt(src.ptr, n, dst.ptr, log_n, inv, shift_mont=fr_to_mont(shift) if shift else None)
    got = _dev_words(zctx, dst.ptr, n).reshape(-1)
    expect = cref.ntt_words(a.reshape(-1), n, log_n, N.omega_for_size(n), inv, shift, threads=0)
    assert np.array_equal(got, expect)
    src.free()
    dst.free()


@pytest.mark.parametrize("pass_bits,tile_bits", [(0, 0), (10, 0), (12, 12)])
def test_ntt_2_20_extreme_inputs(zctx, pass_bits, tile_bits):
    """Buffers may hold any 256-bit Montgomery word, canonical or not.  The
    largest ones (2^256 - 1, r - 1, v + r) with passes of ten stages give the
    largest intermediate values the lazily reduced butterflies see (on the
    2^10- and on the 2^12-element tiles)."""
    log_n = 20
    n = 1 << log_n
    rng = np.random.default_rng(0x2003)
    table = [(1 << 256) - 1, R_MOD - 1]
    for j in range(4094):
        v = int.from_bytes(rng.bytes(32), "little") % R_MOD
        table.append(v + R_MOD if j % 2 == 0 and v + R_MOD < (1 << 256) else v)
    rinv = pow(1 << 256, -1, R_MOD)
    mask = (1 << 64) - 1
    raw_t = np.array([[v >> (64 * j) & mask for j in range(4)] for v in table], dtype=np.uint64)
    plain_t = np.array([[(v * rinv % R_MOD) >> (64 * j) & mask for j in range(4)] for v in table], dtype=np.uint64)
    idx = rng.integers(0, len(table), n)
    idx[0::4], idx[1::4] = 0, 1      # a quarter each of the two largest words
    raw = raw_t[idx].tobytes()
    plain = plain_t[idx].reshape(-1)
    src = zctx.upload(raw)
    dst = zctx.alloc(32 * n)
    zctx.set_tuning("ntt_max_pass_bits", pass_bits)
    zctx.set_tuning("ntt_tile_bits", tile_bits)
    try:
        for inv, shift in ((False, None), (True, FR_GENERATOR)):
            zctx.ntt(src.ptr, n, dst.ptr, log_n, inv, shift_mont=fr_to_mont(shift) if shift else None)
            got = _dev_words(zctx, dst.ptr, n).reshape(-1)
            expect = cref.ntt_words(plain, n, log_n, N.omega_for_size(n), inv, shift, threads=0)
            assert np.array_equal(got, expect), (inv, pass_bits)
    finally:
        zctx.set_tuning("ntt_max_pass_bits", 0)
        zctx.set_tuning("ntt_tile_bits", 0)
    src.free()
    dst.free()


@pytest.mark.parametrize("pass_bits", [0, 12])
def test_ntt_2_24_bit_exact(zctx, pass_bits):
    """2^24 against the C oracle in full (forward and coset inverse), with the
    default three passes of eight stages and with two passes of twelve on
    the 2^12-element tiles."""
    log_n = 24
    n = 1 << log_n
    a = workload.fr_uniform(0x2002, n)
    src = workload.device_fr_mont(zctx, a)
    dst = zctx.alloc(32 * n)
    zctx.set_tuning("ntt_max_pass_bits", pass_bits)
    try:
        for inv, shift in ((False, None), (True, FR_GENERATOR)):
            print(f"2^24 pass_bits={pass_bits} inverse={inv}", flush=True)
            zctx.ntt(src.ptr, n, dst.ptr, log_n, inv, shift_mont=fr_to_mont(shift) if shift else None)
            got = _dev_words(zctx, dst.ptr, n).reshape(-1)
            expect = cref.ntt_words(a.reshape(-1), n, log_n, N.omega_for_size(n), inv, shift, threads=0)
            assert np.array_equal(got, expect), (inv, pass_bits)
    finally:
        zctx.set_tuning("ntt_max_pass_bits", 0)
    src.free()
    dst.free()


def test_ntt_2_24_roundtrip_and_rows(zctx):
    log_n = 24
    n = 1 << log_n
    a = workload.fr_uniform(0x2002, n)
    src = workload.device_fr_mont(zctx, a)
    ev = zctx.alloc(32 * n)
    back = zctx.alloc(32 * n)
    zctx.ntt(src.ptr, n, ev.ptr, log_n, False)
    got = _dev_words(zctx, ev.ptr, n)
    w = N.omega_for_size(n)
    flat = a.reshape(-1)
    for i in (0, 1, 2, 12345, n // 2, n - 1):
        assert words_to_ints(got[i:i + 1])[0] == cref.poly_eval_words(flat, n, pow(w, i, R_MOD)), i
    zctx.ntt(ev.ptr, n, back.ptr, log_n, True)
    assert np.array_equal(_dev_words(zctx, back.ptr, n), a)
    # coset round trip in place
    sh = fr_to_mont(FR_GENERATOR)
    zctx.ntt(src.ptr, n, ev.ptr, log_n, False, shift_mont=sh)
    got = _dev_words(zctx, ev.ptr, n)
    g = FR_GENERATOR
    for i in (0, 7, n - 3):
        assert words_to_ints(got[i:i + 1])[0] == cref.poly_eval_words(flat, n, g * pow(w, i, R_MOD) % R_MOD)
    zctx.ntt(ev.ptr, n, ev.ptr, log_n, True, shift_mont=sh)
    assert np.array_equal(_dev_words(zctx, ev.ptr, n), a)
    for b in (src, ev, back):
        b.free()


def test_ntt_2_28_max_size(zctx):
    """The largest domain Fr has (2-adicity 28): 8 GiB per buffer, so element
    byte offsets pass 2^32.  A sparse input (48 nonzero coefficients) is
    checked at 16 outputs against direct evaluation, plain and on the coset,
    and must come back from the in-place coset inverse with zeros elsewhere;
    a dense uniform input must round-trip bit-exactly."""
    log_n = 28
    n = 1 << log_n
    mont = 1 << 256
    rinv = pow(mont, -1, R_MOD)
    rng = np.random.default_rng(0x2828)
    src = zctx.alloc(32 * n)
    ev = zctx.alloc(32 * n)

    def get(buf, i):
        return int.from_bytes(zctx.download(buf.ptr + 32 * i, 32), "little") * rinv % R_MOD

    pos = sorted(set([0, 1, n // 2, n - 1] + [int(x) for x in rng.integers(0, n, 44)]))
    coef = {p: int.from_bytes(rng.bytes(32), "little") % R_MOD for p in pos}
    zctx.memset_zero(src.ptr, 32 * n)
    staged = [(p, (v * mont % R_MOD).to_bytes(32, "little")) for p, v in coef.items()]
    for p, b in staged:
        zctx.h2d(src.ptr + 32 * p, b, 32)
    zctx.sync()

    def ev_at(x):
        return sum(v * pow(x, p, R_MOD) for p, v in coef.items()) % R_MOD

    w = N.omega_for_size(n)
    outs = [0, 1, 2, 3, n // 4, n // 2, n - 1] + [int(x) for x in rng.integers(0, n, 9)]
    zctx.ntt(src.ptr, n, ev.ptr, log_n, False)
    zctx.sync()
    for i in outs:
        assert get(ev, i) == ev_at(pow(w, i, R_MOD)), i
    sh = fr_to_mont(FR_GENERATOR)
    zctx.ntt(src.ptr, n, ev.ptr, log_n, False, shift_mont=sh)
    zctx.sync()
    for i in outs[:8]:
        assert get(ev, i) == ev_at(FR_GENERATOR * pow(w, i, R_MOD) % R_MOD), i
    zctx.ntt(ev.ptr, n, ev.ptr, log_n, True, shift_mont=sh)
    zctx.sync()
    for p in pos:
        assert get(ev, p) == coef[p], p
    assert zctx.fr_count_nonzero(ev.ptr, n) == len(pos)

    # dense round trip: canonical words (top word below 2^61, so below r)
    a = rng.integers(0, np.iinfo(np.uint64).max, size=(n, 4), dtype=np.uint64, endpoint=True)
    a[:, 3] &= np.uint64((1 << 61) - 1)
    zctx.h2d(src.ptr, a.ctypes.data, 32 * n)
    zctx.sync()
    zctx.fr_to_mont(src.ptr, src.ptr, n)
    zctx.ntt(src.ptr, n, ev.ptr, log_n, False)
    zctx.ntt(ev.ptr, n, ev.ptr, log_n, True)
    zctx.fr_from_mont(ev.ptr, ev.ptr, n)
    back = np.empty((n, 4), dtype="<u8")
    zctx.d2h(back.ctypes.data, ev.ptr, 32 * n)
    zctx.sync()
    assert np.array_equal(back, a)
    src.free()
    ev.free()


def test_fixed_base_kernel(zctx):
    ks = workload.fr_uniform(0x1002, 300)
    ks[5] = 0
    ks[6] = 0
    ks[6, 0] = 1
    bases = workload.device_bases(zctx, ks)
    tmp = zctx.alloc(64 * 300)
    zctx.fq_from_mont(bases.ptr, tmp.ptr, 600)
    xy = words_to_ints(np.frombuffer(zctx.download(tmp.ptr, 64 * 300), dtype="<u8").reshape(-1, 4))
    for i, k in enumerate(words_to_ints(ks)):
        p = None if xy[2 * i] == 0 and xy[2 * i + 1] == 0 else (xy[2 * i], xy[2 * i + 1])
        assert p == C.g1_mul_affine(C.G1_GENERATOR, k), i


def _msm(ctx, bases, scal, n, c=0):
    p = from_jacobian_mont(*ctx.msm_g1(bases.ptr, scal.ptr, n, c))
    return None if p[2] == 0 else (p[0], p[1])


def _closed(ks, ss):
    return M.msm_closed_form(words_to_ints(ks), words_to_ints(ss))


def test_msm_2_12_config(zctx):
    n = 1 << 12
    ks = workload.fr_uniform(0x1002, n)
    ss = workload.fr_uniform(0x1001, n)
    bases = workload.device_bases(zctx, ks)
    scal = workload.device_fr_mont(zctx, ss)
    got = _msm(zctx, bases, scal, n, 10)
    assert got == _closed(ks, ss)
    tmp = zctx.alloc(64 * n)
    zctx.fq_from_mont(bases.ptr, tmp.ptr, 2 * n)
    bw = np.frombuffer(zctx.download(tmp.ptr, 64 * n), dtype="<u8")
    assert cref.msm_pippenger_words(bw, ss.reshape(-1), n, M.default_window_bits(n), threads=0) == got


@pytest.mark.parametrize("log_n,skewed", [(16, False), (16, True), (22, False), (22, True)])
def test_msm_closed_form_large(zctx, log_n, skewed):
    n = 1 << log_n
    ks = workload.fr_uniform(0x3002, n)
    if skewed:
        ss = workload.scalars_skewed(0x3003, n)
        ks = workload.repeat_some(0x3004, ks, 0.05)
    else:
        ss = workload.fr_uniform(0x3001, n)
    bases = workload.device_bases(zctx, ks)
    scal = workload.device_fr_mont(zctx, ss)
    expect = _closed(ks, ss)
    cs = (13, 16, 18) if log_n == 22 else (8, 13, 16)
    for c in cs:
        assert _msm(zctx, bases, scal, n, c) == expect, c
    # the same bases as a prepared (converted once) set: full size and a prefix
    mb = zctx.msm_bases_create(bases.ptr, n)
    for c in cs[:2]:
        p = from_jacobian_mont(*zctx.msm_g1_prepared(mb, scal.ptr, n, c))
        assert (None if p[2] == 0 else (p[0], p[1])) == expect, ("prepared", c)
    m = n // 3 + 1
    p = from_jacobian_mont(*zctx.msm_g1_prepared(mb, scal.ptr, m, 0))
    assert (None if p[2] == 0 else (p[0], p[1])) == _closed(ks[:m], ss[:m])
    with pytest.raises(ValueError):
        zctx.msm_g1_prepared(mb, scal.ptr, n + 1, 0)
    mb.free()
    bases.free()
    scal.free()


def test_msm_2_24_closed_form(zctx):
    # the largest shape here: three-level scans in the bucket sort
    n = 1 << 24
    ks = workload.fr_uniform(0x3102, n)
    ss = workload.fr_uniform(0x3101, n)
    bases = workload.device_bases(zctx, ks)
    scal = workload.device_fr_mont(zctx, ss)
    expect = _closed(ks, ss)
    assert _msm(zctx, bases, scal, n, 17) == expect
    bases.free()
    scal.free()


@pytest.mark.parametrize("c", [0, 8, 13])
def test_fixed_base_batch(zctx, c):
    n = 1 << 12
    ks = workload.fr_uniform(0x4101, n)
    bases = workload.device_bases(zctx, ks)
    fb = zctx.fixed_bases_create(bases.ptr, n, c)
    sizes = [1, 0, 5, 100, 1000, n, 777, 2]
    tot = sum(sizes)
    ss = workload.scalars_skewed(0x4102, tot)
    scal = workload.device_fr_mont(zctx, ss)
    offs = [sum(sizes[:j]) for j in range(len(sizes))]
    got = zctx.msm_g1_fixed_batch(fb, scal.ptr, offs, sizes)
    kk = words_to_ints(ks)
    sv = words_to_ints(ss)
    for j, (o, m) in enumerate(zip(offs, sizes)):
        p = from_jacobian_mont(*got[j])
        res = None if p[2] == 0 else (p[0], p[1])
        assert res == M.msm_closed_form(kk[:m], sv[o:o + m]), (c, j, m)
    fb.free()
    bases.free()
    scal.free()


def test_fixed_base_batch_same_scalars(zctx):
    # every scalar equal: one bucket holds everything (deep fix-up)
    n = 3000
    ks = workload.fr_uniform(0x4201, n)
    bases = workload.device_bases(zctx, ks)
    fb = zctx.fixed_bases_create(bases.ptr, n, 10)
    s = workload.fr_uniform(0x4202, 1)
    ss = np.repeat(s, n, axis=0)
    scal = workload.device_fr_mont(zctx, ss)
    got = zctx.msm_g1_fixed_batch(fb, scal.ptr, [0, 0], [n, n // 2])
    kk = words_to_ints(ks)
    sv = words_to_ints(ss)
    for (o, m), g in zip(((0, n), (0, n // 2)), got):
        p = from_jacobian_mont(*g)
        assert (p[0], p[1]) == M.msm_closed_form(kk[:m], sv[:m])
    fb.free()
