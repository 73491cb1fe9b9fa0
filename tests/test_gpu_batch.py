"""Batch plans: fixed-base MSMs of uneven sizes plus NTT jobs, eager and
replayed from a hipGraph; every replay checked (inputs refilled between
replays, with other eager work in between)."""
import numpy as np
import pytest

from oracle import cref
from oracle import msm as M
from oracle import ntt as N
from pyzkp_amd import workload
from pyzkp_amd.curve import from_jacobian_mont
from pyzkp_amd.field import FR_GENERATOR, fr_to_mont, words_to_ints

pytestmark = pytest.mark.gpu


def _canon(ctx, ptr, n):
    tmp = ctx.alloc(32 * n)
    ctx.fr_from_mont(ptr, tmp.ptr, n)
    w = np.frombuffer(ctx.download(tmp.ptr, 32 * n), dtype="<u8").reshape(n, 4).copy()
    tmp.free()
    return w


def _fill(ctx, dst_ptr, words):
    n = words.shape[0]
    ctx.h2d(dst_ptr, np.ascontiguousarray(words, dtype="<u8").tobytes(), 32 * n)
    ctx.fr_to_mont(dst_ptr, dst_ptr, n)
    ctx.sync()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_plan(zctx, use_graph):
    nb = 1 << 12
    ks = workload.fr_uniform(0x4301, nb)
    bases = workload.device_bases(zctx, ks)
    fb = zctx.fixed_bases_create(bases.ptr, nb, 11)
    sizes = [nb, 1, 300, 0, 2000, 17, 4095, 1024]
    offs = [sum(sizes[:j]) for j in range(len(sizes))]
    tot = sum(sizes)
    scal = zctx.alloc(32 * tot)
    # NTT jobs: two sizes, one batched, one coset inverse
    j1 = dict(log_n=10, batch=3, inverse=False)
    j2 = dict(log_n=13, batch=1, inverse=True, shift_mont=fr_to_mont(FR_GENERATOR))
    bufs = []
    for j in (j1, j2):
        n = (1 << j["log_n"]) * j["batch"]
        j["in_buf"], j["out_buf"] = zctx.alloc(32 * n), zctx.alloc(32 * n)
        j["in"], j["out"] = j["in_buf"].ptr, j["out_buf"].ptr
    plan = zctx.batch_create(fb, scal.ptr, offs, sizes, [j1, j2], use_graph=use_graph)
    kk = words_to_ints(ks)
    for rep in range(3):
        ss = workload.scalars_skewed(0x4400 + rep, tot)
        _fill(zctx, scal.ptr, ss)
        ins = []
        for j in (j1, j2):
            n = (1 << j["log_n"]) * j["batch"]
            a = workload.fr_uniform(0x4500 + rep * 7 + j["log_n"], n)
            _fill(zctx, j["in"], a)
            ins.append(a)
        # unrelated eager work between replays
        tmp = zctx.alloc(32 * 4096)
        zctx.ntt(j1["in"], 1024, tmp.ptr, 12, False)
        zctx.sync()
        tmp.free()
        res = plan.run()
        sv = words_to_ints(ss)
        for jj, (o, m) in enumerate(zip(offs, sizes)):
            p = from_jacobian_mont(*res[jj])
            got = None if p[2] == 0 else (p[0], p[1])
            assert got == M.msm_closed_form(kk[:m], sv[o:o + m]), (rep, jj)
        # NTT outputs
        n1 = 1 << 10
        out1 = _canon(zctx, j1["out"], 3 * n1)
        for b in range(3):
            expect = cref.ntt_words(ins[0][b * n1:(b + 1) * n1].reshape(-1), n1, 10, N.omega_for_size(n1), False)
            assert np.array_equal(out1[b * n1:(b + 1) * n1].reshape(-1), expect)
        n2 = 1 << 13
        out2 = _canon(zctx, j2["out"], n2)
        expect = cref.ntt_words(ins[1].reshape(-1), n2, 13, N.omega_for_size(n2), True, FR_GENERATOR)
        assert np.array_equal(out2.reshape(-1), expect)
    plan.free()
    fb.free()
