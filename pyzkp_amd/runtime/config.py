"""Runtime policy flags (reference: runtime/config.py:10-54), with the
Metal A/B switches replaced by the HIP path's tuning knobs."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Dict

from .context import CPUContext, DeviceContext, HipContext
from .ir.types import Backend


@dataclass(frozen=True)
class RuntimeConfig:
    backend: Backend = Backend.CPU
    cache_dir: str = ".pyZKP_cache"

    fixed_base_policy: str = "off"
    fixed_base_window_bits: int = 8
    fixed_base_auto_min_points: int = 256
    fixed_base_auto_groth16_min_calls: int = 2

    reuse_graph: bool = False
    reuse_prove_batch: bool = False

    # HIP path: MSM window bits (0 = the library default rule / tuned table),
    # NTT stages per pass (0 = planner default), device ordinal
    hip_msm_window_bits: int = 0
    hip_ntt_max_pass_bits: int = 0
    hip_device: int = 0

    def runtime_attrs(self) -> Dict[str, Any]:
        return {
            "fixed_base_policy": str(self.fixed_base_policy),
            "fixed_base_window_bits": int(self.fixed_base_window_bits),
            "fixed_base_auto_min_points": int(self.fixed_base_auto_min_points),
            "fixed_base_auto_groth16_min_calls": int(self.fixed_base_auto_groth16_min_calls),
            "hip_msm_window_bits": int(self.hip_msm_window_bits),
            "hip_ntt_max_pass_bits": int(self.hip_ntt_max_pass_bits),
        }

    def with_overrides(self, overrides: Dict[str, Any] | None) -> Dict[str, Any]:
        out = dict(self.runtime_attrs())
        if overrides:
            out.update(dict(overrides))
        return out

    def make_context(self, *, pool=None, context: DeviceContext | None = None) -> DeviceContext:
        if context is not None:
            return context
        if self.backend == Backend.HIP:
            ctx = HipContext.create_default(pool=pool, config=self, ordinal=self.hip_device)
            if self.hip_ntt_max_pass_bits:
                ctx.hip.set_tuning("ntt_max_pass_bits", self.hip_ntt_max_pass_bits)
            return ctx
        return CPUContext(pool=pool, backend=self.backend, config=self)
