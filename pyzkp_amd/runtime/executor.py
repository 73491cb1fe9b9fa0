"""DAG executor (reference: runtime/executor.py:26-228).

Runs a Graph in topological order.  For each node the target device is the
context's device when the registry has a kernel for (op, device, backend),
else the CPU; inputs that live on the other device get a TO_DEVICE /
FROM_DEVICE transfer inserted on the fly (named "{id}_{device}"), run
through the transfer kernels registered under Device.CPU.  With ``keep``
given, dead intermediates are released to the pool.  With a trace, each
kernel call is timed; HIP kernels are stream-ordered, so the executor waits
for the device before closing a traced event.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Any, Dict

from .config import RuntimeConfig
from .context import CPUContext, DeviceContext
from .ir.graph import Graph, GraphAnalysis, Node
from .ir.ops import OpType
from .ir.types import Backend, Buffer, Device, DType
from .kernels.registry import KernelRegistry
from .memory import MemoryPool
from .trace import Trace, TraceEvent


@dataclass
class Executor:
    registry: KernelRegistry

    def run(self, graph: Graph, *, trace: Trace | None = None, pool: MemoryPool | None = None,
            keep: list[str] | None = None, backend: Backend = Backend.CPU,
            context: DeviceContext | None = None, runtime_config: RuntimeConfig | None = None) -> None:
        analysis = graph.analyze_cached()
        if context is not None:
            ctx = context
        elif runtime_config is not None:
            ctx = runtime_config.make_context(pool=pool)
        else:
            ctx = CPUContext(backend=backend, pool=pool)
        if ctx.pool is None:
            ctx.pool = pool
        self._run_with_analysis(graph, analysis, trace=trace, pool=ctx.pool, keep=keep, backend=ctx.backend,
                                context=ctx)

    def run_repeated(self, graph: Graph, repeat: int, *, trace: Trace | None = None,
                     pool: MemoryPool | None = None, keep: list[str] | None = None,
                     backend: Backend = Backend.CPU, context: DeviceContext | None = None,
                     before_each=None) -> None:
        analysis = graph.analyze_cached()
        ctx = context or CPUContext(backend=backend, pool=pool)
        if ctx.pool is None:
            ctx.pool = pool
        keep0 = set(keep) if keep is not None else set()
        keep0 |= set(analysis.initial)
        for i in range(int(repeat)):
            if before_each is not None:
                before_each(i, graph)
            self._run_with_analysis(graph, analysis, trace=trace, pool=ctx.pool, keep=list(keep0),
                                    backend=ctx.backend, context=ctx)

    # ------------------------------------------------------------------ core
    def _release(self, graph: Graph, bid: str, pool: MemoryPool | None) -> None:
        buf = graph.buffers.get(bid)
        if buf is None:
            return
        if pool is not None:
            if buf.device == Device.CPU and buf.dtype == DType.FR and isinstance(buf.data, list):
                pool.release_cpu(DType.FR, buf.data)
            elif buf.device == Device.HIP and hasattr(buf.data, "dbuf"):
                pool.release_device(buf.data.dbuf)
        del graph.buffers[bid]

    def _run_with_analysis(self, graph: Graph, analysis: GraphAnalysis, *, trace: Trace | None,
                           pool: MemoryPool | None, keep: list[str] | None, backend: Backend,
                           context: DeviceContext) -> None:
        keep_set = set(keep) if keep is not None else None
        use_count: Dict[str, int] = {}
        for node in graph.nodes:
            for inp in node.inputs:
                use_count[inp] = use_count.get(inp, 0) + 1

        for idx in analysis.topo_order:
            node = graph.nodes[idx]
            target = Device.CPU
            if backend != Backend.CPU and self.registry.has(node.op, context.device, backend=backend):
                target = context.device
            new_inputs = []
            for inp_id in node.inputs:
                buf = graph.buffers[inp_id]
                if buf.device != target:
                    op = OpType.TO_DEVICE if target != Device.CPU else OpType.FROM_DEVICE
                    tid = f"{inp_id}_{target.value}"
                    if tid not in graph.buffers:
                        tnode = Node(op=op, inputs=[inp_id], outputs=[tid], attrs={})
                        self._run_node(graph, tnode, trace=trace, pool=pool, backend=backend, context=context,
                                       force_device=Device.CPU)
                    use_count[tid] = use_count.get(tid, 0) + 1
                    new_inputs.append(tid)
                    if keep_set is not None:
                        use_count[inp_id] = use_count.get(inp_id, 0) - 1
                        if use_count[inp_id] == 0 and inp_id not in keep_set:
                            self._release(graph, inp_id, pool)
                else:
                    new_inputs.append(inp_id)
            if new_inputs != node.inputs:
                node = Node(op=node.op, inputs=new_inputs, outputs=node.outputs, attrs=node.attrs)
                graph.nodes[idx] = node
            self._run_node(graph, node, trace=trace, pool=pool, backend=backend, context=context,
                           force_device=target)
            if keep_set is not None:
                for inp in node.inputs:
                    use_count[inp] = use_count.get(inp, 0) - 1
                    if use_count[inp] == 0 and inp not in keep_set:
                        self._release(graph, inp, pool)

    def _run_node(self, graph: Graph, node: Node, *, trace: Trace | None, pool: MemoryPool | None,
                  backend: Backend, context: DeviceContext, force_device: Device | None = None) -> None:
        if len(node.inputs) == 0:
            raise ValueError("node must have inputs")
        device = force_device if force_device is not None else graph.buffers[node.inputs[0]].device
        fn = self.registry.get(node.op, device, backend=backend)
        inputs = [graph.buffers[i] for i in node.inputs]
        ctx: Dict[str, Any] = {"graph": graph, "node": node, "inputs": inputs, "attrs": dict(node.attrs),
                               "pool": pool, "context": context}
        in_sizes = [_buffer_size(b) for b in inputs]
        t0 = time.perf_counter_ns()
        out = fn(ctx)
        if trace is not None and getattr(context, "hip", None) is not None:
            context.hip.sync()
        t1 = time.perf_counter_ns()
        if "outputs" not in out:
            raise ValueError("kernel must return outputs")
        outputs: Dict[str, Buffer] = out["outputs"]
        for bid, buf in outputs.items():
            graph.buffers[bid] = buf
        if trace is not None:
            trace.add(TraceEvent(op=node.op, device=device, backend=backend, start_ns=t0, end_ns=t1,
                                 attrs=dict(node.attrs), input_sizes=in_sizes,
                                 output_sizes=[_buffer_size(outputs[k]) for k in node.outputs if k in outputs]))


def _buffer_size(buf: Buffer) -> int:
    d = buf.data
    if isinstance(d, (list, tuple, bytes, bytearray)):
        return len(d)
    if hasattr(d, "n"):
        return int(d.n)
    return 1
