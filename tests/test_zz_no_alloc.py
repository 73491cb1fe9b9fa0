"""zk_msm_g1_noalloc — the MSM on the library's file-scope device arena,
with no device allocation — through the C-ABI, against zk_msm_g1 on the
same inputs and against the closed form [sum s_i k_i] G for bases [k_i] G:
two calls in a row, the second with a smaller input, then a shape whose
arena would not fit, refused with ZK_ERR_INVALID."""
import ctypes as ct

import pytest

from oracle import msm as M
from pyzkp_amd import workload
from pyzkp_amd.curve import from_jacobian_mont
from pyzkp_amd.field import words_to_ints

pytestmark = pytest.mark.gpu


def _affine(words):
    f = lambda a: words[a] | (words[a + 1] << 64) | (words[a + 2] << 128) | (words[a + 3] << 192)
    p = from_jacobian_mont(f(0), f(4), f(8))
    return None if p[2] == 0 else (p[0], p[1])


def _noalloc(ctx, bases, scal, n, c):
    out = (ct.c_uint64 * 12)()
    rc = ctx.lib.zk_msm_g1_noalloc(ctx.handle, bases.ptr, scal.ptr, n, c, out, None)
    return rc, list(out)


def _inputs(ctx, n, seed):
    ks = workload.fr_uniform(seed, n)
    ss = workload.fr_uniform(seed + 1, n)
    return ks, ss, workload.device_bases(ctx, ks), workload.device_fr_mont(ctx, ss)


@pytest.mark.parametrize("shapes", [[(1 << 15, 0), (5000, 0)], [(1 << 16, 14), (1 << 12, 10)]])
def test_noalloc_matches_msm_g1(zctx, shapes):
    for n, c in shapes:
        print(f"noalloc MSM n={n} c={c}", flush=True)
        ks, ss, bases, scal = _inputs(zctx, n, 0x7000 + n)
        rc, out = _noalloc(zctx, bases, scal, n, c)
        assert rc == 0, zctx.lib.zk_last_error(zctx.handle)
        got = _affine(out)
        ref = from_jacobian_mont(*zctx.msm_g1(bases.ptr, scal.ptr, n, c))
        assert got == (None if ref[2] == 0 else (ref[0], ref[1]))
        assert got == M.msm_closed_form(words_to_ints(ks), words_to_ints(ss))
        bases.free()
        scal.free()


def test_noalloc_skewed_and_repeated(zctx):
    """Skewed scalars (zeros, small values) and repeated bases: large buckets
    spanning many segments, and the doubling case, on the static arena."""
    n = 1 << 14
    ks = workload.repeat_some(0x7204, workload.fr_uniform(0x7202, n), 0.05)
    ss = workload.scalars_skewed(0x7203, n)
    bases, scal = workload.device_bases(zctx, ks), workload.device_fr_mont(zctx, ss)
    for c in (8, 13):
        rc, out = _noalloc(zctx, bases, scal, n, c)
        assert rc == 0
        assert _affine(out) == M.msm_closed_form(words_to_ints(ks), words_to_ints(ss)), c
    bases.free()
    scal.free()


def test_noalloc_refuses_oversized_shape(zctx):
    n = 1 << 17
    ks, ss, bases, scal = _inputs(zctx, n, 0x7100)
    rc, out = _noalloc(zctx, bases, scal, n, 0)
    assert rc == 1
    assert b"scratch bytes" in zctx.lib.zk_last_error(zctx.handle)
    assert out == [0] * 12
    # a fitting call right after still works
    rc, out = _noalloc(zctx, bases, scal, 1 << 10, 0)
    assert rc == 0
    assert _affine(out) == M.msm_closed_form(words_to_ints(ks[:1 << 10]), words_to_ints(ss[:1 << 10]))
    bases.free()
    scal.free()
