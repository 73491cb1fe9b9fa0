"""Autotuner for the HIP path: picks the MSM window size, the MSM
accumulation segment length, the fixed-base table window and the NTT pass
size by timing candidates on the device, and caches the winners per
(device, shape) in a JSON file (``PYZKP_AMD_TUNE_CACHE``, default
``~/.cache/pyzkp_amd/tune.json``).  The reference has no counterpart (its
window rule is fixed, runtime/kernels/cpu/kernels.py:557-560).
"""
from __future__ import annotations

import json
import os
import time
from typing import Callable, Dict, Iterable, Optional, Tuple

_DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "pyzkp_amd", "tune.json")


def _time(fn: Callable[[], object], sync: Callable[[], None], reps: int = 2) -> float:
    fn()
    sync()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        sync()
        best = min(best, time.perf_counter() - t0)
    return best


class Tuner:
    def __init__(self, ctx, cache_path: Optional[str] = None, use_cache: bool = True) -> None:
        self.ctx = ctx
        self.path = cache_path or os.environ.get("PYZKP_AMD_TUNE_CACHE", _DEFAULT_CACHE)
        self.use_cache = use_cache
        self.cache: Dict[str, dict] = {}
        if use_cache and os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    self.cache = json.load(f)
            except (OSError, ValueError):
                self.cache = {}
        self.log: Dict[str, Dict[str, float]] = {}

    def _save(self) -> None:
        if not self.use_cache:
            return
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.cache, f, indent=1)
        except OSError:
            pass

    def _pick(self, key: str, cands: Iterable, run: Callable[[object], float]) -> object:
        if self.use_cache and key in self.cache:
            return self.cache[key]["best"]
        times = {}
        for cnd in cands:
            times[cnd] = run(cnd)
        best = min(times, key=times.get)
        self.log[key] = {str(k): round(v * 1e3, 4) for k, v in times.items()}
        self.cache[key] = {"best": best, "ms": self.log[key]}
        self._save()
        return best

    # ---- plain MSM: window size c
    def msm_window(self, bases: int, scalars: int, n: int, cands: Iterable[int] = range(13, 19)) -> int:
        ctx = self.ctx
        return self._pick(f"msm_c/n={n}", list(cands),
                          lambda c: _time(lambda: ctx.msm_g1(bases, scalars, n, c), ctx.sync))

    # ---- plain MSM: accumulation segment length
    def msm_segment(self, bases: int, scalars: int, n: int, c: int,
                    cands: Iterable[int] = (0, 32, 48, 64, 80, 96, 128)) -> int:
        ctx = self.ctx

        def run(seg):
            ctx.set_tuning("msm_segment", seg)
            return _time(lambda: ctx.msm_g1(bases, scalars, n, c), ctx.sync)

        best = self._pick(f"msm_segment/n={n}/c={c}", list(cands), run)
        ctx.set_tuning("msm_segment", best)
        return best

    # ---- fixed-base batch: table window c (builds one table per candidate)
    def fixed_window(self, bases: int, nb: int, scalars: int, offsets, sizes,
                     cands: Iterable[int] = range(10, 18)) -> Tuple[int, object]:
        ctx = self.ctx
        key = f"fixed_c/nb={nb}/count={len(sizes)}/total={sum(sizes)}"

        def run(c):
            fb = ctx.fixed_bases_create(bases, nb, c)
            t = _time(lambda: ctx.msm_g1_fixed_batch(fb, scalars, offsets, sizes), ctx.sync)
            fb.free()
            return t

        best = self._pick(key, list(cands), run)
        return best, ctx.fixed_bases_create(bases, nb, best)

    # ---- NTT: stages per pass
    def ntt_pass_bits(self, src: int, dst: int, log_n: int, cands: Iterable[int] = (6, 7, 8, 9, 10, 11)) -> int:
        ctx = self.ctx
        n = 1 << log_n

        def run(bits):
            ctx.set_tuning("ntt_max_pass_bits", bits)
            return _time(lambda: ctx.ntt(src, n, dst, log_n, False), ctx.sync)

        best = self._pick(f"ntt_pass_bits/log_n={log_n}", list(cands), run)
        ctx.set_tuning("ntt_max_pass_bits", best)
        return best
