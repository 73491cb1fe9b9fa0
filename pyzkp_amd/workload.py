"""Synthetic inputs for tests and benchmarks (vectorised with numpy).

The generator is splitmix64; an Fr element is 4 successive words w0..w3
(little endian) with w3 &= 2^62 - 1, accepted if < r, else redrawn from the
next 4 words.  A uniform float is (next() >> 11) / 2^53.  Inputs are
returned as (n, 4) uint64 arrays of canonical values.
"""
from __future__ import annotations

import numpy as np

from .field import FR_MODULUS

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)
_R_WORDS = [(FR_MODULUS >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)]


def splitmix_block(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs number start+1 .. start+count of the splitmix64 stream."""
    k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
    return z ^ (z >> np.uint64(31))


def _less_than_r(w: np.ndarray) -> np.ndarray:
    lt = np.zeros(w.shape[0], dtype=bool)
    eq = np.ones(w.shape[0], dtype=bool)
    for i in (3, 2, 1, 0):
        ri = np.uint64(_R_WORDS[i])
        lt |= eq & (w[:, i] < ri)
        eq &= w[:, i] == ri
    return lt


def fr_uniform(seed: int, n: int) -> np.ndarray:
    """n canonical Fr elements as (n, 4) uint64."""
    out = np.empty((0, 4), dtype=np.uint64)
    drawn = 0  # candidates consumed so far
    while out.shape[0] < n:
        need = n - out.shape[0]
        m = int(need * 1.4) + 64
        raw = splitmix_block(seed, 4 * drawn, 4 * m).reshape(m, 4)
        raw[:, 3] &= np.uint64((1 << 62) - 1)
        ok = _less_than_r(raw)
        acc = raw[ok]
        if acc.shape[0] >= need:
            # drop candidates after the n-th acceptance so later blocks stay aligned
            acc = acc[:need]
        out = np.concatenate([out, acc])
        drawn += m
    return out


def uniform01(seed: int, n: int) -> np.ndarray:
    return (splitmix_block(seed, 0, n) >> np.uint64(11)).astype(np.float64) / float(1 << 53)


# ---------------------------------------------------------------- MSM inputs
def scalars_skewed(seed: int, n: int) -> np.ndarray:
    """Bucket-imbalance stress: per i with u ~ U[0,1) (stream seed):
    u < 0.10 -> 0, u < 0.30 -> uniform below 2^16, else uniform in Fr."""
    u = uniform01(seed, n)
    full = fr_uniform(seed ^ 0x5EED, n)
    small = splitmix_block(seed ^ 0x5111, 0, n) & np.uint64(0xFFFF)
    out = full.copy()
    z = u < 0.10
    s = (u >= 0.10) & (u < 0.30)
    out[z] = 0
    out[s] = 0
    out[s, 0] = small[s]
    return out


def repeat_some(seed: int, ks: np.ndarray, frac: float = 0.05) -> np.ndarray:
    """With probability frac, k_i := k_j for a uniformly random earlier j."""
    n = ks.shape[0]
    u = uniform01(seed, n)
    pick = (splitmix_block(seed ^ 0xB00B, 0, n) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    out = ks.copy()
    for i in np.nonzero(u < frac)[0]:
        if i == 0:
            continue
        j = int(pick[i] * i)
        out[i] = out[j]
    return out


_TABLE_CACHE: dict = {}


def device_fr_mont(ctx, words: np.ndarray):
    """(n, 4) canonical uint64 words -> device Montgomery Fr buffer."""
    n = words.shape[0]
    buf = ctx.alloc(max(32 * n, 32))
    if n:
        ctx.h2d(buf.ptr, np.ascontiguousarray(words, dtype="<u8").tobytes(), 32 * n)
        ctx.fr_to_mont(buf.ptr, buf.ptr, n)
        ctx.sync()
    return buf


def device_bases(ctx, ks_words: np.ndarray):
    """Affine Montgomery bases B_i = [k_i] G1 built on the device."""
    from .curve import fixed_base_table

    n = ks_words.shape[0]
    key = ctx.handle.value if hasattr(ctx.handle, "value") else id(ctx)
    tab = _TABLE_CACHE.get(key)
    if tab is None:
        tab = ctx.upload(fixed_base_table())
        _TABLE_CACHE[key] = tab
    kbuf = device_fr_mont(ctx, ks_words)
    out = ctx.alloc(max(64 * n, 64))
    ctx.g1_fixed_base(tab.ptr, kbuf.ptr, n, out.ptr)
    ctx.sync()
    kbuf.free()
    return out


# ---------------------------------------------------------------- batched MSM + NTT workload
def batch_workload(ctx, srs_log: int = 18, count: int = 64, log_lo: int = 10, ntt_lo: int = 12,
                   ntt_hi: int = 18) -> dict:
    """The batched workload: `count` MSMs with n_j = floor(2^U) points, U ~
    U[log_lo, srs_log] (seed 0x4001), over one shared 2^srs_log-point SRS
    (bases [k_i] G1, k seed 0x4003; scalars seed 0x4004, MSM j reads scalars
    [offs_j, offs_j + n_j)), plus `count` forward NTTs with log2 size uniform
    in ntt_lo..ntt_hi (seed 0x4002), grouped by size into batched jobs (the
    inputs of size 2^lg from seed 0x4005 + lg).  Device buffers are in
    `buffers` (the caller frees them)."""
    nb = 1 << srs_log
    u = uniform01(0x4001, count)
    sizes = [int(2.0 ** (log_lo + (srs_log - log_lo) * x)) for x in u]
    offs = [int(v) for v in np.cumsum([0] + sizes[:-1])]
    ks = fr_uniform(0x4003, nb)
    ss = fr_uniform(0x4004, sum(sizes))
    bases = device_bases(ctx, ks)
    scal = device_fr_mont(ctx, ss)
    span = ntt_hi - ntt_lo + 1
    logs = [ntt_lo + int(v) % span for v in splitmix_block(0x4002, 0, count)]
    jobs, inputs, bufs = [], {}, [bases, scal]
    for lg in sorted(set(logs)):
        cnt = logs.count(lg)
        a = fr_uniform(0x4005 + lg, cnt << lg)
        src = device_fr_mont(ctx, a)
        dst = ctx.alloc(32 * (cnt << lg))
        bufs += [src, dst]
        inputs[lg] = a
        jobs.append({"in": src.ptr, "out": dst.ptr, "log_n": lg, "batch": cnt, "inverse": False})
    return {"srs_n": nb, "ks": ks, "ss": ss, "sizes": sizes, "offs": offs, "bases": bases, "scalars": scal,
            "ntt_logs": logs, "ntt_jobs": jobs, "ntt_inputs": inputs, "buffers": bufs}
