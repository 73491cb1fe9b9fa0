"""Deterministic input generator shared by tests and bench (oracle copy).

splitmix64 stream; an Fr element is 4 successive words w0..w3 (little
endian) with w3 &= 2^62 - 1, accepted if < r, else the next 4 words are
drawn.  A uniform float is (next() >> 11) / 2^53.  The product keeps its own
vectorised copy (pyzkp_amd/workload.py); tests check the two agree.
"""
from __future__ import annotations

from typing import List

from .field import R_MOD

_M64 = (1 << 64) - 1
_GAMMA = 0x9E3779B97F4A7C15


class SplitMix64:
    def __init__(self, seed: int) -> None:
        self.state = seed & _M64

    def next(self) -> int:
        self.state = (self.state + _GAMMA) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) / float(1 << 53)


def fr_uniform(rng: SplitMix64) -> int:
    while True:
        w = [rng.next() for _ in range(4)]
        w[3] &= (1 << 62) - 1
        x = w[0] | (w[1] << 64) | (w[2] << 128) | (w[3] << 192)
        if x < R_MOD:
            return x


def fr_vector(seed: int, n: int) -> List[int]:
    rng = SplitMix64(seed)
    return [fr_uniform(rng) for _ in range(n)]
