"""Continuous-batching Groth16 prove server on one GPU.

Proving keys of several circuits stay resident on the device (query
fixed-base tables, R1CS matrices); each device slot keeps its own NTT tables.
Requests (circuit, witness, r, s) arrive over time.

  * Admission.  A request is checked on arrival (known circuit, witness
    shape, working set within the HBM budget); its G2 MSM is queued at once
    on a host pool (shortest job first), since it needs nothing from the
    device.
  * Device slots.  The server owns `device_slots` contexts on the GPU, each
    with its own HIP stream and scratch arenas and one host thread.  A slot
    runs one group at a time: up to `max_batch` requests of ONE circuit,
    whose H pipelines run back to back and whose G1 MSMs run as one
    fixed-base launch.  Different circuits' groups run on different slots
    at the same time, so their device work overlaps.
  * When a group is cut.  The G2 MSM on the host is the longest stage of a
    proof, so a request's device work is not on the critical path until its
    G2 MSM starts.  The pending requests of a circuit therefore form a group
    when the first of them reaches a G2 worker, or when max_batch of them
    wait, and a slot and HBM are free: under light load every group holds
    one request and starts as soon as its G2 MSM does; under load requests
    wait for G2 workers anyway, and the device batches them meanwhile.
  * HBM accounting.  The working sets of the groups on the slots (scalar
    segments plus the per-group a, b, c and scratch vectors) never exceed
    the budget; a group waits for memory to free up.
  * Completion.  When a request's G1 parts (device) and G2 part (host) are
    both ready, a small assembler pool forms the proof.
  * Device G2 (g2_device).  The G2 MSMs run on the group's slot after its
    G1 batch, no host G2 workers are used, and a group forms whenever a
    slot and HBM are free, so the batch size follows the load.

A witness that fails its R1CS fails only its own request (its group is
re-run one request at a time).  `serve_plain` is the baseline: `prove()`
once per request in arrival order.  The reference has no serving layer;
its closest pattern is prove_batch (protocols/groth16/prove.py:133-248),
sequential re-use of one key.
"""
from __future__ import annotations

import heapq
import importlib
import queue
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

from .groth16.types import Proof, ProvingKey

# the prover module itself (the package re-exports a function of the same name)
P = importlib.import_module(".groth16.prove", __package__)


@dataclass
class Request:
    id: int
    key: object                 # circuit id (key of the server's proving-key dict)
    witness: np.ndarray         # (n_vars, 4) canonical words
    r: int
    s: int
    arrival: float = 0.0        # seconds after the run starts


@dataclass
class Result:
    id: int
    proof: Optional[Proof]
    arrival: float
    admitted: float = 0.0       # when its device group started
    done: float = 0.0
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _HostPool:
    """Host workers for the G2 MSMs, shortest job first: a queued job with
    fewer bases runs before a larger one, so a small proof does not wait for
    the G2 MSMs of large proofs that arrived before it."""

    def __init__(self, workers: int) -> None:
        self._heap: list = []
        self._cv = threading.Condition()
        self._seq = 0
        self._stop = False
        self._threads = [threading.Thread(target=self._loop, daemon=True) for _ in range(workers)]
        for t in self._threads:
            t.start()

    def submit(self, size: int, fn, *args) -> Future:
        fut: Future = Future()
        with self._cv:
            heapq.heappush(self._heap, (size, self._seq, fut, fn, args))
            self._seq += 1
            self._cv.notify()
        return fut

    def _loop(self) -> None:
        while True:
            with self._cv:
                while not self._heap and not self._stop:
                    self._cv.wait()
                if not self._heap:
                    return
                _, _, fut, fn, args = heapq.heappop(self._heap)
            try:
                fut.set_result(fn(*args))
            except BaseException as e:     # handed to the waiter
                fut.set_exception(e)

    def shutdown(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()


def working_set_bytes(pk: ProvingKey) -> int:
    """Device bytes one in-flight proof of this key needs: its scalar segment
    [w | h | l] plus the per-group a, b, c and scratch vectors (counted per
    request, an upper bound)."""
    return P.segment_bytes(pk) + 32 * 4 * pk.n


class _Slot:
    """One device context (own stream, arenas and NTT tables) with its host
    thread and its per-circuit workspaces."""

    def __init__(self, index: int, ctx, owns_ctx: bool) -> None:
        self.index = index
        self.ctx = ctx
        self.owns_ctx = owns_ctx
        self.work: Dict[object, object] = {}    # circuit -> prover workspace on this context
        self._q: "queue.Queue" = queue.Queue()
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    def workspace(self, key, pk: ProvingKey, k: int):
        self.work[key] = P.grow_work(pk, self.work.get(key), k, self.ctx)
        return self.work[key]

    def workspace_bytes(self) -> int:
        return sum(w.nbytes for w in self.work.values())

    def submit(self, fn, *args) -> None:
        self._q.put((fn, args))

    def _loop(self) -> None:
        while True:
            job = self._q.get()
            if job is None:
                return
            fn, args = job
            try:
                fn(self, *args)
            except BaseException:       # jobs record their own errors; the slot must keep serving
                pass

    def close(self) -> None:
        self._q.put(None)
        self._thread.join()
        self.ctx.sync()
        for w in self.work.values():
            w.free()
        self.work.clear()
        if self.owns_ctx:
            self.ctx.close()


class ProveServer:
    def __init__(self, keys: Dict[object, ProvingKey], hbm_budget: int, max_batch: int = 16,
                 g2_workers: int = 2, cpu_threads: int = 16, device_slots: int = 2,
                 assemblers: int = 2, slot_contexts: Optional[list] = None, g2_device: bool = False) -> None:
        """`slot_contexts`: contexts for the device slots (owned by the
        caller); by default each slot opens a context of its own on the
        keys' device.  `g2_device`: each group's G2 MSMs run on its slot
        after the G1 batch instead of on the host G2 workers; a group then
        forms as soon as a slot is free (every pending request of the first
        circuit, up to max_batch)."""
        if device_slots < 1 or max_batch < 1:
            raise ValueError("device_slots and max_batch must be at least 1")
        self.keys = keys
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.cpu_threads = cpu_threads
        self.g2_device = bool(g2_device)
        self.slots: List[_Slot] = []
        if slot_contexts is not None:
            self.slots = [_Slot(i, c, False) for i, c in enumerate(slot_contexts)]
        else:
            ords = {pk.ctx.ordinal for pk in keys.values()}
            if len(ords) != 1:
                raise ValueError("the proving keys must all live on one device")
            from ._native import Context
            ordinal = next(iter(ords))
            self.slots = [_Slot(i, Context(ordinal), True) for i in range(device_slots)]
        if not self.slots:
            raise ValueError("the server needs at least one device slot")
        self.pool = _HostPool(g2_workers)
        self.assembler = ThreadPoolExecutor(max_workers=assemblers)
        self._reset_stats()

    def _reset_stats(self) -> None:
        self.stats = {"groups": 0, "batches": 0, "msm_launches": 0, "max_batch_seen": 0,
                      "max_concurrent_groups": 0, "peak_inflight_bytes": 0, "device_slots": len(self.slots),
                      "workspace_bytes": sum(s.workspace_bytes() for s in self.slots)}

    def close(self) -> None:
        self.pool.shutdown()
        self.assembler.shutdown(wait=True)
        for s in self.slots:
            s.close()
        self.slots = []

    def footprint(self, req: Request) -> int:
        return working_set_bytes(self.keys[req.key])

    # ------------------------------------------------------------------ run
    def run(self, requests: List[Request]) -> List[Result]:
        """Feeds `requests` at their arrival times, serves them and returns
        the results in request order.  `stats` describes this run; the
        slots (contexts, NTT tables, workspaces) persist across runs."""
        reqs = sorted(requests, key=lambda q: q.arrival)
        for q in reqs:
            if q.key not in self.keys:
                raise ValueError(f"request {q.id}: unknown circuit {q.key!r}")
            if self.footprint(q) > self.budget:
                raise ValueError(f"request {q.id} needs {self.footprint(q)} device bytes, above the HBM budget "
                                 f"{self.budget}")
        self._reset_stats()
        out = _Run(self, reqs).serve(requests)
        self.stats["workspace_bytes"] = sum(s.workspace_bytes() for s in self.slots)
        return out


class _Run:
    """State of one ProveServer.run call."""

    def __init__(self, srv: ProveServer, reqs: List[Request]) -> None:
        self.srv = srv
        self.reqs = reqs
        self.cv = threading.Condition()
        self.pending: List[Request] = []        # arrived, device group not formed yet
        self.g2_started: set = set()            # ids whose G2 MSM is running or done
        self.free_slots = list(srv.slots)
        self.inflight = 0                       # working-set bytes of the groups on the slots
        self.busy = 0
        self.arrived = 0
        self.results: Dict[int, Result] = {}
        self.parts: Dict[int, dict] = {}        # id -> {"g1": ..., "g2": ..., "err": ...}
        self.admitted: Dict[int, float] = {}
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    # ---- arrival and host stages
    def _feed(self) -> None:
        for q in self.reqs:
            delay = q.arrival - self.now()
            if delay > 0:
                time.sleep(delay)
            self._arrive(q)

    def _arrive(self, q: Request) -> None:
        pk = self.srv.keys[q.key]
        try:
            P.check_witness(pk, q.witness)
        except ValueError as e:
            with self.cv:
                self.arrived += 1
                self._finish(q, None, str(e))
                self.cv.notify_all()
            return
        with self.cv:
            self.parts[q.id] = {}
            self.pending.append(q)
            self.arrived += 1
            self.cv.notify_all()
        if not self.srv.g2_device:
            self.srv.pool.submit(pk.n_vars, self._g2_job, q)

    def _g2_job(self, q: Request) -> None:
        with self.cv:
            self.g2_started.add(q.id)
            self.cv.notify_all()
        try:
            out = ("g2", P.g2_part(self.srv.keys[q.key], q.witness, self.srv.cpu_threads))
        except Exception as e:                  # recorded per request
            out = ("err", f"{type(e).__name__}: {e}")
        self._deliver(q, *out)

    def _deliver(self, q: Request, kind: str, value) -> None:
        with self.cv:
            part = self.parts[q.id]
            part[kind] = value
            complete = "err" in part or ("g1" in part and "g2" in part)
            if complete and not part.get("assembling"):
                part["assembling"] = True
            else:
                return
        self.srv.assembler.submit(self._assemble, q)

    def _assemble(self, q: Request) -> None:
        part = self.parts[q.id]
        proof, err = None, part.get("err")
        if err is None:
            try:
                proof = P.assemble(self.srv.keys[q.key], part["g1"], part["g2"], q.r, q.s)
            except Exception as e:
                err = f"{type(e).__name__}: {e}"
        with self.cv:
            self._finish(q, proof, err)
            self.cv.notify_all()

    def _finish(self, q: Request, proof, err) -> None:
        """Records q's result (lock held).  A request that failed before its
        device group formed leaves the pending list, so no device work is
        started for it."""
        self.results[q.id] = Result(q.id, proof, q.arrival, self.admitted.get(q.id, 0.0), self.now(), err)
        if proof is None:
            self.pending = [p for p in self.pending if p.id != q.id]

    # ---- device groups
    def _ready_group(self) -> Optional[List[Request]]:
        """The next group to start, or None (called with the lock held)."""
        if not self.free_slots or not self.pending:
            return None
        srv = self.srv
        by_key: Dict[object, List[Request]] = {}
        for q in self.pending:                  # arrival order within a circuit
            by_key.setdefault(q.key, []).append(q)
        for key, qs in sorted(by_key.items(), key=lambda kv: kv[1][0].arrival):
            due = srv.g2_device or len(qs) >= srv.max_batch or any(q.id in self.g2_started for q in qs)
            if not due:
                continue
            group, used = [], self.inflight
            for q in qs[:srv.max_batch]:
                fp = srv.footprint(q)
                if used + fp > srv.budget:
                    break
                group.append(q)
                used += fp
            if group:
                return group
        return None

    def _launch(self, group: List[Request]) -> None:
        """Starts `group` on a free slot (lock held)."""
        srv = self.srv
        slot = self.free_slots.pop()
        ids = {q.id for q in group}
        self.pending = [q for q in self.pending if q.id not in ids]
        nbytes = sum(srv.footprint(q) for q in group)
        self.inflight += nbytes
        self.busy += 1
        t = self.now()
        for q in group:
            self.admitted[q.id] = t
        st = srv.stats
        st["groups"] += 1
        st["batches"] = st["groups"]
        st["max_batch_seen"] = max(st["max_batch_seen"], len(group))
        st["max_concurrent_groups"] = max(st["max_concurrent_groups"], self.busy)
        st["peak_inflight_bytes"] = max(st["peak_inflight_bytes"], self.inflight)
        slot.submit(self._device_group, group, nbytes)

    def _device_group(self, slot: _Slot, group: List[Request], nbytes: int) -> None:
        """Runs on the slot's thread: the group's device work on its context."""
        key = group[0].key
        pk = self.srv.keys[key]
        outcome: Dict[int, list] = {}
        launches = 0
        g2 = self.srv.g2_device

        def parts_of(part) -> list:
            return [("g1", part[0]), ("g2", part[1])] if g2 else [("g1", part)]

        try:
            try:
                parts = P.device_batch(pk, [q.witness for q in group], slot.workspace(key, pk, len(group)), g2=g2)
                launches += 1
                outcome = {q.id: parts_of(part) for q, part in zip(group, parts)}
            except ValueError:
                # a bad witness fails the group's check: isolate it
                for q in group:
                    try:
                        part = P.device_batch(pk, [q.witness], slot.workspace(key, pk, 1), g2=g2)[0]
                        outcome[q.id] = parts_of(part)
                        launches += 1
                    except ValueError as e:
                        outcome[q.id] = [("err", str(e))]
        except Exception as e:                  # device failure: every request of the group fails
            outcome = {q.id: [("err", f"{type(e).__name__}: {e}")] for q in group}
        with self.cv:
            self.srv.stats["msm_launches"] += launches
            self.inflight -= nbytes
            self.busy -= 1
            self.free_slots.append(slot)
            self.cv.notify_all()
        for q in group:
            for kind, value in outcome[q.id]:
                self._deliver(q, kind, value)

    def serve(self, requests: List[Request]) -> List[Result]:
        feeder = threading.Thread(target=self._feed, daemon=True)
        feeder.start()
        total = len(self.reqs)
        with self.cv:
            # done when every request has its result and no group is still on
            # a slot (so no device work of this run outlives it)
            while len(self.results) < total or self.busy:
                group = self._ready_group()
                if group is not None:
                    self._launch(group)
                    continue
                self.cv.wait(timeout=0.5)
        feeder.join()
        return [self.results[q.id] for q in requests]


def serve_plain(keys: Dict[object, ProvingKey], requests: List[Request], cpu_threads: int = 16,
                g2_device: bool = False) -> List[Result]:
    """Baseline: one proof at a time in arrival order, each with prove() (its
    G2 MSM on `cpu_threads` host threads, overlapped only with its own
    device work, or on the device with g2_device)."""
    t0 = time.perf_counter()
    out = {}
    for q in sorted(requests, key=lambda x: x.arrival):
        delay = q.arrival - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        start = time.perf_counter() - t0
        try:
            pf = P.prove(keys[q.key], q.witness, q.r, q.s, g2_threads=cpu_threads, g2_device=g2_device)
            out[q.id] = Result(q.id, pf, q.arrival, start, time.perf_counter() - t0)
        except ValueError as e:
            out[q.id] = Result(q.id, None, q.arrival, start, time.perf_counter() - t0, str(e))
    return [out[q.id] for q in requests]


def latency_stats(results: List[Result]) -> dict:
    ok = [r for r in results if r.proof is not None]
    if not ok:
        return {"count": 0}
    lat = np.array([r.latency for r in ok])
    span = max(r.done for r in ok) - min(r.arrival for r in ok)
    return {"count": len(ok), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "mean_ms": round(float(lat.mean()) * 1e3, 3), "proofs_per_s": round(len(ok) / span, 3)}


def request_classes(n_req: int, log_lo: int, log_hi: int, seed: int = 0x5002) -> List[int]:
    """Padded circuit size class per request: m = floor(2^U) constraints with
    U ~ U[log_lo, log_hi] (seeded), padded to 2^k >= m as compile_r1cs pads
    (protocols/groth16/r1cs.py:43)."""
    from . import workload
    out = []
    for u in workload.uniform01(seed, n_req):
        m = int(2.0 ** (log_lo + (log_hi - log_lo) * float(u)))
        out.append(max(log_lo, (m - 1).bit_length()))
    return out


def poisson_arrivals(n_req: int, rate: float, seed: int = 0x5001) -> List[float]:
    """Arrival times (seconds) of a Poisson process of `rate` per second."""
    from . import workload
    u = workload.uniform01(seed, n_req)
    gaps = -np.log1p(-np.minimum(u, 1.0 - 2.0 ** -53)) / rate
    return [float(x) for x in np.cumsum(gaps)]


def make_requests(circuits: Dict[int, object], classes: List[int], arrivals: List[float],
                  witness_seed: int = 0x5100, rs_seed: int = 0x5200) -> List[Request]:
    """Request i proves circuit classes[i] with witness seed witness_seed + i
    and (r, s) drawn from seed rs_seed + i."""
    from . import workload
    from .field import words_to_ints
    from .groth16.synthetic import synthetic_witness
    reqs = []
    for i, (k, t) in enumerate(zip(classes, arrivals)):
        w = synthetic_witness(circuits[k], witness_seed + i)
        r, s = words_to_ints(workload.fr_uniform(rs_seed + i, 2))
        reqs.append(Request(id=i, key=k, witness=w, r=r, s=s, arrival=t))
    return reqs
