"""Operator graph: buffers + nodes, static checks and Kahn topological order
(reference: runtime/ir/graph.py:22-153, same API and error behaviour)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Dict, List, Sequence, Tuple

from .ops import OpType
from .types import Buffer, Device, DType


@dataclass(frozen=True)
class Node:
    op: OpType
    inputs: List[str]
    outputs: List[str]
    attrs: Dict[str, Any]


@dataclass(frozen=True)
class GraphAnalysis:
    producers: Dict[str, int]
    topo_order: Tuple[int, ...]
    initial: Tuple[str, ...]


class Graph:
    def __init__(self) -> None:
        self.buffers: Dict[str, Buffer] = {}
        self.nodes: List[Node] = []
        self.meta: Dict[str, Any] = {}
        self._version = 0
        self._analysis_cache: GraphAnalysis | None = None
        self._analysis_cache_version = -1

    def add_buffer(self, *, id: str, device: Device, dtype: DType, data: Any,
                   meta: Dict[str, Any] | None = None) -> Buffer:
        if id in self.buffers:
            raise ValueError(f"buffer already exists: {id}")
        buf = Buffer(id=id, device=device, dtype=dtype, data=data, meta=meta)
        self.buffers[id] = buf
        self._version += 1
        self._analysis_cache = None
        return buf

    def add_node(self, *, op: OpType, inputs: Sequence[str], outputs: Sequence[str],
                 attrs: Dict[str, Any] | None = None) -> Node:
        n = Node(op=op, inputs=list(inputs), outputs=list(outputs), attrs=attrs or {})
        self.nodes.append(n)
        self._version += 1
        self._analysis_cache = None
        return n

    def analyze(self) -> GraphAnalysis:
        producers: Dict[str, int] = {}
        initial = set(self.buffers.keys())
        for idx, node in enumerate(self.nodes):
            if len(node.outputs) == 0:
                raise ValueError(f"node has no outputs: {idx}")
            for out in node.outputs:
                if out in initial:
                    raise ValueError(f"node output overwrites existing buffer: {out}")
                if out in producers:
                    raise ValueError(f"buffer produced by multiple nodes: {out}")
                producers[out] = idx
        indeg = [0] * len(self.nodes)
        succ: List[List[int]] = [[] for _ in self.nodes]
        for j, node in enumerate(self.nodes):
            if len(node.inputs) == 0:
                raise ValueError(f"node has no inputs: {j}")
            for inp in node.inputs:
                if inp in producers:
                    succ[producers[inp]].append(j)
                    indeg[j] += 1
                elif inp not in initial:
                    raise ValueError(f"missing input buffer: {inp}")
        q = [i for i, d in enumerate(indeg) if d == 0]
        topo: List[int] = []
        while q:
            i = q.pop()
            topo.append(i)
            for j in succ[i]:
                indeg[j] -= 1
                if indeg[j] == 0:
                    q.append(j)
        if len(topo) != len(self.nodes):
            raise ValueError("graph has cycle or unresolved dependencies")
        return GraphAnalysis(producers=producers, topo_order=tuple(topo), initial=tuple(sorted(initial)))

    def analyze_cached(self) -> GraphAnalysis:
        if self._analysis_cache is not None and self._analysis_cache_version == self._version:
            return self._analysis_cache
        a = self.analyze()
        self._analysis_cache = a
        self._analysis_cache_version = self._version
        return a

    def reset_to_initial(self, analysis: GraphAnalysis) -> None:
        keep = set(analysis.initial)
        for k in list(self.buffers.keys()):
            if k not in keep:
                del self.buffers[k]
