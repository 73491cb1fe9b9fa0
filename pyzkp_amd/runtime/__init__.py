"""Operator runtime mirroring the reference's (runtime/ in pyZKP): OpType /
Buffer / Graph IR, KernelRegistry, DAG Executor with automatic device
transfers, MemoryPool, RuntimeConfig and Trace — with a HIP backend."""
from .config import RuntimeConfig
from .context import CPUContext, DeviceContext, HipContext
from .executor import Executor
from .ir import Backend, Buffer, Device, DType, Graph, GraphAnalysis, Node, OpType
from .kernels import DeviceBases, HipBuffer, KernelRegistry, register_hip_kernels
from .memory import CPUMemoryPool, MemoryPool, PoolStats
from .trace import Trace, TraceEvent, export_chrome_tracing

__all__ = [
    "RuntimeConfig", "CPUContext", "DeviceContext", "HipContext", "Executor", "Backend", "Buffer", "Device",
    "DType", "Graph", "GraphAnalysis", "Node", "OpType", "DeviceBases", "HipBuffer", "KernelRegistry",
    "register_hip_kernels", "CPUMemoryPool", "MemoryPool", "PoolStats", "Trace", "TraceEvent",
    "export_chrome_tracing",
]
