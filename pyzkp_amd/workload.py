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
