"""pyzkp_amd — MI355X-native prover hot path for a pyZKP-style stack.

BN254 G1 MSM and Fr NTT / iNTT / coset NTT as hand-written HIP kernels for
gfx950 (``csrc/``), reached through a ctypes C-ABI (``include/zkhip.h``,
``_native``), and registered as a ``hip`` backend in a KernelRegistry /
Executor runtime that mirrors the reference's operator API (``runtime``).
"""
from ._native import LIB_PATH, Context, DeviceBuffer, ZkError, get_context, native_available

__all__ = ["LIB_PATH", "Context", "DeviceBuffer", "ZkError", "get_context", "native_available"]
