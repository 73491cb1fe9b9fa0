"""Per-op trace events, aggregation and Chrome-trace export (reference:
runtime/trace.py:18-163)."""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional

from .ir.ops import OpType
from .ir.types import Backend, Device


@dataclass
class TraceEvent:
    op: OpType
    device: Device
    backend: Backend
    start_ns: int
    end_ns: int
    attrs: Dict[str, Any] = field(default_factory=dict)
    input_sizes: List[int] = field(default_factory=list)
    output_sizes: List[int] = field(default_factory=list)

    @property
    def duration_ns(self) -> int:
        return int(self.end_ns - self.start_ns)


class Trace:
    def __init__(self) -> None:
        self.events: List[TraceEvent] = []

    def add(self, ev: TraceEvent) -> None:
        self.events.append(ev)

    def clear(self) -> None:
        self.events.clear()

    def total_ns(self) -> int:
        return sum(e.duration_ns for e in self.events)

    def by_op(self) -> Dict[str, Dict[str, float]]:
        out: Dict[str, Dict[str, float]] = {}
        for e in self.events:
            k = e.op.value if isinstance(e.op, OpType) else str(e.op)
            d = out.setdefault(k, {"count": 0, "total_ms": 0.0})
            d["count"] += 1
            d["total_ms"] += e.duration_ns / 1e6
        return out

    def by_device(self) -> Dict[str, float]:
        out: Dict[str, float] = {}
        for e in self.events:
            out[e.device.value] = out.get(e.device.value, 0.0) + e.duration_ns / 1e6
        return out


def _jsonable(v: Any) -> Any:
    if isinstance(v, (int, float, str, bool)) or v is None:
        return v
    return str(v)


def export_chrome_tracing(trace: Trace, path: str, *, pid: int = 1) -> None:
    """One swimlane (tid) per device, complete ('X') events in microseconds."""
    if not trace.events:
        events: List[Dict[str, Any]] = []
    else:
        t0 = min(e.start_ns for e in trace.events)
        tids = {Device.CPU: 1, Device.HIP: 2}
        events = [{"name": "thread_name", "ph": "M", "pid": pid, "tid": tid, "args": {"name": dev.value}}
                  for dev, tid in tids.items()]
        for e in trace.events:
            events.append({
                "name": e.op.value if isinstance(e.op, OpType) else str(e.op), "ph": "X", "pid": pid,
                "tid": tids.get(e.device, 3), "ts": (e.start_ns - t0) / 1e3, "dur": e.duration_ns / 1e3,
                "args": {"backend": e.backend.value, "inputs": e.input_sizes, "outputs": e.output_sizes,
                         **{k: _jsonable(v) for k, v in e.attrs.items()}},
            })
    with open(path, "w") as f:
        json.dump({"traceEvents": events, "displayTimeUnit": "ms"}, f)


tracer: Optional[Trace] = Trace()
