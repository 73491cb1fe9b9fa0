from .graph import Graph, GraphAnalysis, Node
from .ops import OpType
from .types import Backend, Buffer, Device, DType

__all__ = ["Graph", "GraphAnalysis", "Node", "OpType", "Backend", "Buffer", "Device", "DType"]
