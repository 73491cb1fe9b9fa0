from .registry import KernelFn, KernelRegistry
from .hip import DeviceBases, HipBuffer, register_hip_kernels

__all__ = ["KernelFn", "KernelRegistry", "DeviceBases", "HipBuffer", "register_hip_kernels"]
