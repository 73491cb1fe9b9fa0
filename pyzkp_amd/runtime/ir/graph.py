"""Operator graph: named buffers and operator nodes.

Public surface kept from the reference (runtime/ir/graph.py:22-153) so that
graphs built against it run unchanged: `Node`, `GraphAnalysis` (producers,
topo_order, initial), `Graph.add_buffer` / `add_node` / `analyze` /
`analyze_cached` / `reset_to_initial`, and ValueError for every structural
fault (duplicate buffer, output redefining a buffer, missing input, node
without inputs or outputs, cycle).

Design here:
  * Sources are tracked explicitly.  A buffer is a source when it was added
    with `add_buffer`, or when it sits in `buffers` without any node
    producing it and is not a device copy the executor made.  Buffers
    written by an earlier run are therefore not mistaken for sources, and
    `analyze()` gives the same answer before and after a run.
  * The order is the smallest-index-first topological order: among the
    nodes whose inputs are ready, the one added first runs first.  A graph
    written in dependency order runs exactly in insertion order, which
    makes traces predictable; out-of-order insertion is still accepted.
  * A cycle is reported with the positions of the nodes it leaves
    unscheduled.
"""
from __future__ import annotations

import heapq
from dataclasses import dataclass
from typing import Any, Dict, List, Optional, Sequence, Set, Tuple

from .ops import OpType
from .types import Buffer, Device, DType


@dataclass(frozen=True)
class Node:
    """One operator application: `op` reads the buffers named in `inputs`
    and defines the buffers named in `outputs`; `attrs` holds its
    parameters (sizes, roots, shifts, challenges)."""
    op: OpType
    inputs: List[str]
    outputs: List[str]
    attrs: Dict[str, Any]


@dataclass(frozen=True)
class GraphAnalysis:
    """Result of `Graph.analyze`: which node defines each buffer
    (`producers`, node positions in `Graph.nodes`), a valid execution order
    (`topo_order`) and the source buffer ids (`initial`, sorted)."""
    producers: Dict[str, int]
    topo_order: Tuple[int, ...]
    initial: Tuple[str, ...]


class Graph:
    def __init__(self) -> None:
        self.buffers: Dict[str, Buffer] = {}
        self.nodes: List[Node] = []
        self.meta: Dict[str, Any] = {}
        self._declared: Set[str] = set()          # ids given to add_buffer
        self._derived: Set[str] = set()           # device copies inserted by the executor
        self._cached: Optional[GraphAnalysis] = None

    # ------------------------------------------------------------ building
    def add_buffer(self, *, id: str, device: Device, dtype: DType, data: Any,
                   meta: Dict[str, Any] | None = None) -> Buffer:
        if id in self.buffers or id in self._declared:
            raise ValueError(f"buffer id {id!r} is already defined in this graph")
        buf = Buffer(id=id, device=device, dtype=dtype, data=data, meta=meta)
        self.buffers[id] = buf
        self._declared.add(id)
        self._cached = None
        return buf

    def add_node(self, *, op: OpType, inputs: Sequence[str], outputs: Sequence[str],
                 attrs: Dict[str, Any] | None = None) -> Node:
        node = Node(op=op, inputs=list(inputs), outputs=list(outputs), attrs=dict(attrs) if attrs else {})
        self.nodes.append(node)
        self._cached = None
        return node

    # ------------------------------------------------------------ analysis
    def _definitions(self) -> Dict[str, int]:
        """Buffer id -> position of the node that defines it, with the
        single-definition rules checked."""
        defined_by: Dict[str, int] = {}
        for pos, node in enumerate(self.nodes):
            if not node.outputs:
                raise ValueError(f"node {pos} ({node.op.value}) defines no output buffer")
            if not node.inputs:
                raise ValueError(f"node {pos} ({node.op.value}) reads no input buffer")
            for bid in node.outputs:
                if bid in self._declared:
                    raise ValueError(f"node {pos} ({node.op.value}) would redefine source buffer {bid!r}")
                if bid in defined_by:
                    raise ValueError(f"buffer {bid!r} is defined by node {defined_by[bid]} and node {pos}")
                defined_by[bid] = pos
        return defined_by

    def _sources(self, defined_by: Dict[str, int]) -> Set[str]:
        return self._declared | {b for b in self.buffers if b not in defined_by and b not in self._derived}

    def note_derived(self, bid: str) -> None:
        """Marks `bid` as a copy made while running (a device transfer), never
        a source of the graph."""
        self._derived.add(bid)

    def analyze(self) -> GraphAnalysis:
        defined_by = self._definitions()
        sources = self._sources(defined_by)
        waiting = [0] * len(self.nodes)               # unresolved inputs per node
        readers: Dict[int, List[int]] = {}            # producer position -> reader positions
        for pos, node in enumerate(self.nodes):
            for bid in node.inputs:
                src = defined_by.get(bid)
                if src is not None:
                    readers.setdefault(src, []).append(pos)
                    waiting[pos] += 1
                elif bid not in sources:
                    raise ValueError(f"node {pos} ({node.op.value}) reads buffer {bid!r}, which nothing defines")
        ready = [pos for pos, w in enumerate(waiting) if w == 0]
        heapq.heapify(ready)
        order: List[int] = []
        while ready:
            pos = heapq.heappop(ready)
            order.append(pos)
            for nxt in readers.get(pos, ()):
                waiting[nxt] -= 1
                if waiting[nxt] == 0:
                    heapq.heappush(ready, nxt)
        if len(order) < len(self.nodes):
            stuck = sorted(set(range(len(self.nodes))) - set(order))
            raise ValueError(f"graph has a dependency cycle through nodes {stuck}")
        return GraphAnalysis(producers=defined_by, topo_order=tuple(order), initial=tuple(sorted(sources)))

    def analyze_cached(self) -> GraphAnalysis:
        """`analyze()`, recomputed only after the graph's structure changed."""
        if self._cached is None:
            self._cached = self.analyze()
        return self._cached

    def reset_to_initial(self, analysis: GraphAnalysis) -> None:
        """Drops every buffer that is not a source of `analysis`, so the graph
        can run again from its inputs."""
        sources = set(analysis.initial)
        for bid in [b for b in self.buffers if b not in sources]:
            del self.buffers[bid]
