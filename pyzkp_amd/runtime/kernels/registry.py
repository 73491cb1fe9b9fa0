"""(OpType, Device, Backend) -> KernelFn map (reference:
runtime/kernels/registry.py:17-53): duplicate registration raises
ValueError, a missing kernel raises KeyError."""
from __future__ import annotations

from typing import Any, Callable, Dict, Tuple

from ..ir.ops import OpType
from ..ir.types import Backend, Device

KernelFn = Callable[[Dict[str, Any]], Dict[str, Any]]


class KernelRegistry:
    def __init__(self) -> None:
        self._kernels: Dict[Tuple[OpType, Device, Backend], KernelFn] = {}

    def register(self, op: OpType, device: Device, fn: KernelFn, *, backend: Backend = Backend.CPU) -> None:
        key = (op, device, backend)
        if key in self._kernels:
            raise ValueError(f"kernel already registered: {op} on {device} ({backend})")
        self._kernels[key] = fn

    def get(self, op: OpType, device: Device, *, backend: Backend = Backend.CPU) -> KernelFn:
        key = (op, device, backend)
        if key not in self._kernels:
            raise KeyError(f"kernel not found: {op} on {device} ({backend})")
        return self._kernels[key]

    def has(self, op: OpType, device: Device, *, backend: Backend = Backend.CPU) -> bool:
        return (op, device, backend) in self._kernels
