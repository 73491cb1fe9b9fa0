"""Execution contexts (reference: runtime/context.py:11-36): backend,
device, pool and config, plus the HIP device handle for the hip backend."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from .ir.types import Backend, Device


@dataclass
class DeviceContext:
    backend: Backend = Backend.CPU
    device: Device = Device.CPU
    pool: Any | None = None
    config: Any | None = None


@dataclass
class CPUContext(DeviceContext):
    backend: Backend = Backend.CPU
    device: Device = Device.CPU


@dataclass
class HipContext(DeviceContext):
    backend: Backend = Backend.HIP
    device: Device = Device.HIP
    hip: Any | None = None  # pyzkp_amd.Context

    @staticmethod
    def create_default(*, pool=None, config=None, ordinal: int = 0) -> "HipContext":
        from .._native import get_context, native_available

        if not native_available():
            raise RuntimeError("HIP runtime not available (libzkhip.so not built or no HIP device)")
        return HipContext(pool=pool, hip=get_context(ordinal), config=config)
