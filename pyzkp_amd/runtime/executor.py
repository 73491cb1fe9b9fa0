"""Graph executor.

Public surface kept from the reference (runtime/executor.py:31-218):
`Executor(registry).run(graph, trace=, pool=, keep=, backend=, context=,
runtime_config=)` and `run_repeated(graph, repeat, ..., before_each=)`;
the KernelFn call contract (a dict with graph, node, inputs, attrs, pool,
context; the kernel returns {"outputs": {id: Buffer}}, else ValueError);
placement on the context's device whenever the registry has a kernel for
(op, device, backend), else on the CPU; device copies named
"{id}_{device}" made by the TO_DEVICE / FROM_DEVICE kernels registered
under Device.CPU; with `keep`, every buffer that is dead and not kept is
handed back to the pool and removed from the graph.

Design here:
  * Lifetimes come from a liveness pass over the order before anything
    runs: each buffer expires at the step of its last reader, and its
    device copies expire with it.  No per-buffer counters are updated while
    running.
  * The graph's nodes are never rewritten.  A node whose inputs needed a
    device copy is called with a per-call view of itself that names the
    copies, so the same graph can run any number of times.
  * With a trace, each kernel call is timed to completion: HIP kernels are
    stream-ordered, so the device is synchronised before the event closes.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Any, Dict, Iterable, List, Optional, Set

from .config import RuntimeConfig
from .context import CPUContext, DeviceContext
from .ir.graph import Graph, GraphAnalysis, Node
from .ir.ops import OpType
from .ir.types import Backend, Buffer, Device, DType
from .kernels.registry import KernelRegistry
from .memory import MemoryPool
from .trace import Trace, TraceEvent


def _extent(buf: Buffer) -> int:
    """Size of a buffer for the trace: element count where it has one."""
    data = buf.data
    if isinstance(data, (list, tuple, bytes, bytearray)):
        return len(data)
    n = getattr(data, "n", None)
    return int(n) if n is not None else 1


def _recycle(buf: Buffer, pool: Optional[MemoryPool]) -> None:
    """Hands a dead buffer's storage back to the pool it may have come from
    (the pool ignores storage it does not own)."""
    if pool is None:
        return
    if buf.device == Device.CPU and buf.dtype == DType.FR and isinstance(buf.data, list):
        pool.release_cpu(DType.FR, buf.data)
    elif buf.device == Device.HIP and hasattr(buf.data, "dbuf"):
        pool.release_device(buf.data.dbuf)


def _expiry(graph: Graph, order: Iterable[int]) -> Dict[int, List[str]]:
    """Step index -> ids of the buffers whose last reader runs at that step."""
    last: Dict[str, int] = {}
    for step, pos in enumerate(order):
        for bid in graph.nodes[pos].inputs:
            last[bid] = step
    out: Dict[int, List[str]] = {}
    for bid, step in last.items():
        out.setdefault(step, []).append(bid)
    return out


@dataclass
class Executor:
    registry: KernelRegistry

    def run(self, graph: Graph, *, trace: Trace | None = None, pool: MemoryPool | None = None,
            keep: list[str] | None = None, backend: Backend = Backend.CPU,
            context: DeviceContext | None = None, runtime_config: RuntimeConfig | None = None) -> None:
        if context is None:
            context = (runtime_config.make_context(pool=pool) if runtime_config is not None
                       else CPUContext(backend=backend, pool=pool))
        if context.pool is None:
            context.pool = pool
        self._execute(graph, graph.analyze_cached(), context, trace,
                      None if keep is None else set(keep))

    def run_repeated(self, graph: Graph, repeat: int, *, trace: Trace | None = None,
                     pool: MemoryPool | None = None, keep: list[str] | None = None,
                     backend: Backend = Backend.CPU, context: DeviceContext | None = None,
                     before_each=None) -> None:
        """Runs the graph `repeat` times; the graph's sources always survive,
        so every repetition starts from the same inputs.  `before_each(i,
        graph)` runs before repetition i."""
        analysis = graph.analyze_cached()
        if context is None:
            context = CPUContext(backend=backend, pool=pool)
        if context.pool is None:
            context.pool = pool
        retained = set(keep or ()) | set(analysis.initial)
        for i in range(int(repeat)):
            if before_each is not None:
                before_each(i, graph)
            self._execute(graph, analysis, context, trace, retained)

    # ---------------------------------------------------------------- internals
    def _placement(self, op: OpType, context: DeviceContext) -> Device:
        if context.backend != Backend.CPU and self.registry.has(op, context.device, backend=context.backend):
            return context.device
        return Device.CPU

    def _execute(self, graph: Graph, analysis: GraphAnalysis, context: DeviceContext,
                 trace: Optional[Trace], retained: Optional[Set[str]]) -> None:
        order = analysis.topo_order
        expiring = _expiry(graph, order) if retained is not None else {}
        for step, pos in enumerate(order):
            node = graph.nodes[pos]
            where = self._placement(node.op, context)
            args = [self._resident(graph, bid, where, context, trace) for bid in node.inputs]
            view = node if args == node.inputs else Node(op=node.op, inputs=args, outputs=node.outputs,
                                                          attrs=node.attrs)
            self._invoke(graph, view, where, context, trace)
            for bid in expiring.get(step, ()):
                self._expire(graph, bid, retained, context.pool)

    def _resident(self, graph: Graph, bid: str, where: Device, context: DeviceContext,
                  trace: Optional[Trace]) -> str:
        """Id of a copy of buffer `bid` on `where`, made on first need."""
        if graph.buffers[bid].device == where:
            return bid
        copy_id = f"{bid}_{where.value}"
        if copy_id not in graph.buffers:
            op = OpType.FROM_DEVICE if where == Device.CPU else OpType.TO_DEVICE
            self._invoke(graph, Node(op=op, inputs=[bid], outputs=[copy_id], attrs={}), Device.CPU, context, trace)
            graph.note_derived(copy_id)
        return copy_id

    def _expire(self, graph: Graph, bid: str, retained: Set[str], pool: Optional[MemoryPool]) -> None:
        for victim in [bid] + [f"{bid}_{d.value}" for d in Device]:
            if victim in retained:
                continue
            buf = graph.buffers.pop(victim, None)
            if buf is not None:
                _recycle(buf, pool)

    def _invoke(self, graph: Graph, node: Node, where: Device, context: DeviceContext,
                trace: Optional[Trace]) -> None:
        if not node.inputs:
            raise ValueError(f"{node.op.value} node has no inputs")
        kernel = self.registry.get(node.op, where, backend=context.backend)
        inputs = [graph.buffers[b] for b in node.inputs]
        call: Dict[str, Any] = {"graph": graph, "node": node, "inputs": inputs, "attrs": dict(node.attrs),
                                "pool": context.pool, "context": context}
        start = time.perf_counter_ns()
        result = kernel(call)
        if trace is not None and getattr(context, "hip", None) is not None:
            context.hip.sync()
        stop = time.perf_counter_ns()
        produced = result.get("outputs") if isinstance(result, dict) else None
        if produced is None:
            raise ValueError(f"kernel for {node.op.value} returned no outputs")
        graph.buffers.update(produced)
        if trace is not None:
            trace.add(TraceEvent(op=node.op, device=where, backend=context.backend, start_ns=start, end_ns=stop,
                                 attrs=dict(node.attrs), input_sizes=[_extent(b) for b in inputs],
                                 output_sizes=[_extent(produced[b]) for b in node.outputs if b in produced]))
