"""Buffer pool: host FR lists keyed by (dtype, length) and device
allocations keyed by byte size, plus one reusable pinned staging area for
host <-> device copies.

Counterpart of the reference's MemoryPool (runtime/memory.py:20-110),
which pools Python lists and Metal shared (UMA) buffers.  On MI355X device
memory is separate from host memory, so device buffers are plain HBM
allocations reused by size, and copies go through pinned host memory.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Dict, List, Tuple

from .ir.types import DType


@dataclass
class PoolStats:
    alloc_calls: int = 0
    reuse_calls: int = 0
    in_use: int = 0
    peak_in_use: int = 0

    def _take(self, reused: bool) -> None:
        self.alloc_calls += 1
        if reused:
            self.reuse_calls += 1
        self.in_use += 1
        self.peak_in_use = max(self.peak_in_use, self.in_use)


class MemoryPool:
    def __init__(self) -> None:
        self._free_cpu: Dict[Tuple[DType, int], List[List[int]]] = {}
        self._owned_cpu_ids: set = set()
        self.cpu_stats = PoolStats()
        self._free_dev: Dict[Tuple[int, int], List[Any]] = {}  # (ctx id, bytes) -> DeviceBuffers
        self._owned_dev_ids: set = set()
        self.device_stats = PoolStats()
        self._staging: Dict[int, Tuple[int, int]] = {}  # ctx id -> (pinned ptr, bytes)
        self._staging_ctx: Dict[int, Any] = {}          # ctx id -> ctx (to free the area)

    # ---- host lists (reference runtime/memory.py:34-71)
    def alloc_fr(self, n: int) -> List[int]:
        return self.alloc_cpu(DType.FR, n)

    def alloc_cpu(self, dtype: DType, n: int) -> List[int]:
        lst = self._free_cpu.get((dtype, int(n)))
        if lst:
            self.cpu_stats._take(True)
            out = lst.pop()
            for i in range(len(out)):
                out[i] = 0
        else:
            self.cpu_stats._take(False)
            out = [0] * int(n)
        self._owned_cpu_ids.add(id(out))
        return out

    def release_cpu(self, dtype: DType, arr: List[int]) -> None:
        if id(arr) not in self._owned_cpu_ids:
            return
        self._owned_cpu_ids.discard(id(arr))
        self.cpu_stats.in_use -= 1
        self._free_cpu.setdefault((dtype, len(arr)), []).append(arr)

    # ---- device memory (reference runtime/memory.py:75-110, UMA -> HBM)
    def alloc_device(self, ctx, size_bytes: int):
        key = (id(ctx), int(size_bytes))
        lst = self._free_dev.get(key)
        if lst:
            self.device_stats._take(True)
            out = lst.pop()
        else:
            self.device_stats._take(False)
            out = ctx.alloc(int(size_bytes))
        self._owned_dev_ids.add(id(out))
        return out

    def release_device(self, dbuf) -> None:
        if id(dbuf) not in self._owned_dev_ids:
            return
        self._owned_dev_ids.discard(id(dbuf))
        self.device_stats.in_use -= 1
        self._free_dev.setdefault((id(dbuf.ctx), dbuf.nbytes), []).append(dbuf)

    # ---- pinned staging for host <-> device copies
    def staging(self, ctx, nbytes: int) -> int:
        cur = self._staging.get(id(ctx))
        if cur is None or cur[1] < nbytes:
            if cur is not None:
                ctx.sync()
                ctx.pinned_free(cur[0])
            size = max(int(nbytes), 1 << 20)
            ptr = ctx.pinned_alloc(size)
            self._staging[id(ctx)] = (ptr, size)
            self._staging_ctx[id(ctx)] = ctx
            return ptr
        return cur[0]

    def close(self) -> None:
        """Frees the pooled device buffers and the pinned staging areas."""
        for lst in self._free_dev.values():
            for b in lst:
                b.free()
        self._free_dev.clear()
        for ctx_id, (ptr, _) in list(self._staging.items()):
            ctx = self._staging_ctx.pop(ctx_id)
            ctx.sync()
            ctx.pinned_free(ptr)
        self._staging.clear()


# reference alias (runtime/memory.py:112-113)
CPUMemoryPool = MemoryPool
