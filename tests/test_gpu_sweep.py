"""Seeded sweeps over MSM and NTT shapes on the GPU, against the closed form
and the C oracle.

MSM: random sizes up to 2^17 and window sizes 2..24, with scalars that mix
uniform values, zeros, r - 1 and small values, and bases with repeats, run
through the digit-split and the chunked bucket reductions, the fused and the
chunked fine level of the bucket sort, one or several window groups, and
plain and prepared base sets; every result equals [sum s_i k_i] G for bases
[k_i] G.

NTT: random sizes 2^1..2^18, input lengths (zero-padded and truncated
forward transforms), shifts (none, the generator, a random one, zero) and
pass limits, all four variants against the C restatement of the reference
NTT."""
import random

import numpy as np
import pytest

from oracle import cref
from oracle import msm as M
from oracle import ntt as N
from oracle.field import R_MOD
from pyzkp_amd import workload
from pyzkp_amd.curve import from_jacobian_mont
from pyzkp_amd.field import words_to_ints

pytestmark = pytest.mark.gpu


def _words(xs):
    return np.array([[(x >> (64 * i)) & ((1 << 64) - 1) for i in range(4)] for x in xs], dtype=np.uint64)


def _aff(r):
    p = from_jacobian_mont(*r)
    return None if p[2] == 0 else (p[0], p[1])


def test_msm_shape_sweep(zctx):
    rng = random.Random(0x51)
    frng = random.Random(0x53)  # the sort form, drawn apart so the cases above stay as they were
    try:
        for case in range(30):
            n = rng.choice([1, 2, 7, 100, 1000, 4097]) if case < 6 else rng.randrange(1, 1 << 17)
            c = rng.randrange(2, 25)
            split = rng.randrange(2)
            groups = rng.choice([0, 0, 2, 3])
            prepared = rng.randrange(2)
            fused = frng.randrange(2)
            print(f"MSM sweep case {case}: n={n} c={c} split={split} groups={groups} prepared={prepared} "
                  f"fused={fused}", flush=True)
            kw = workload.repeat_some(0x5200 + case, workload.fr_uniform(0x5100 + case, n), 0.1)
            sw = workload.fr_uniform(0x5300 + case, n)
            u = workload.uniform01(0x5400 + case, n)
            sv = words_to_ints(sw)
            for i in range(n):
                if u[i] < 0.1:
                    sv[i] = 0
                elif u[i] < 0.15:
                    sv[i] = R_MOD - 1
                elif u[i] < 0.3:
                    sv[i] = sv[i] & 0xFFFF
            expect = M.msm_closed_form(words_to_ints(kw), sv)
            bases = workload.device_bases(zctx, kw)
            sc = workload.device_fr_mont(zctx, _words(sv))
            zctx.set_tuning("msm_reduce_split", split)
            zctx.set_tuning("msm_groups", groups)
            zctx.set_tuning("msm_sort_fused", fused)
            if prepared:
                mb = zctx.msm_bases_create(bases.ptr, n)
                got = _aff(zctx.msm_g1_prepared(mb, sc.ptr, n, c))
                mb.free()
            else:
                got = _aff(zctx.msm_g1(bases.ptr, sc.ptr, n, c))
            assert got == expect, (case, n, c, split, groups, prepared, fused)
            bases.free()
            sc.free()
    finally:
        zctx.set_tuning("msm_reduce_split", 1)
        zctx.set_tuning("msm_groups", 0)
        zctx.set_tuning("msm_sort_fused", 1)


def test_ntt_shape_sweep(zctx):
    rng = random.Random(0x52)
    try:
        for case in range(48):
            log_n = rng.randrange(1, 19)
            n = 1 << log_n
            inv = bool(rng.randrange(2))
            in_len = n if inv else rng.choice([n, n, max(1, n // 3), n + rng.randrange(1, 50), 1])
            shift = rng.choice([None, None, 5, rng.randrange(1, R_MOD), 0])
            bits = rng.choice([0, 0, rng.randrange(1, 13)])
            print(f"NTT sweep case {case}: n=2^{log_n} inv={inv} in_len={in_len} shift={shift} bits={bits}", flush=True)
            a = [rng.randrange(R_MOD) for _ in range(in_len)]
            src = workload.device_fr_mont(zctx, _words(a))
            dst = zctx.alloc(32 * n)
            zctx.set_tuning("ntt_max_pass_bits", bits)
            zctx.ntt(src.ptr, in_len, dst.ptr, log_n, inv,
                     shift_mont=(shift * (1 << 256) % R_MOD) if shift is not None else None)
            tmp = zctx.alloc(32 * n)
            zctx.fr_from_mont(dst.ptr, tmp.ptr, n)
            zctx.sync()
            got = np.frombuffer(zctx.download(tmp.ptr, 32 * n), dtype="<u8")
            expect = cref.ntt_words(_words(a).reshape(-1), in_len, log_n, N.omega_for_size(n), inv, shift, threads=0)
            assert np.array_equal(got, expect), (case, log_n, inv, in_len, shift, bits)
            for b in (src, dst, tmp):
                b.free()
    finally:
        zctx.set_tuning("ntt_max_pass_bits", 0)
