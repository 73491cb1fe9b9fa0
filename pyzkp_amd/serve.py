"""Continuous-batching Groth16 prove server on one GPU.

Proving keys of several circuits stay resident on the device (query
fixed-base tables, R1CS matrices, NTT tables).  Requests (circuit, witness,
r, s) arrive over time; the scheduler loop

  * admits queued requests in arrival order while the device working sets
    of the batch fit the HBM budget and a batch slot is free (a request
    whose own working set exceeds the budget is rejected with ValueError
    before the run starts);
  * queues each admitted request's G2 MSM on a host pool at once (shortest
    job first; each MSM uses all host threads, the workers share the cores);
  * groups the batch by circuit and runs, per group, the device H
    pipelines back to back and ONE fixed-base launch for all their G1
    MSMs;
  * hands each request's G1 results to its pool task, which assembles the
    proof as soon as its G2 MSM is done, while the scheduler already
    admits the next batch.

So the device work of new requests overlaps the host G2 work of earlier
ones, and a small request is never queued behind a large one's G2 MSM.
`serve_plain` is the baseline: one proof at a time in arrival order.  The
reference has no serving layer; its closest pattern is prove_batch
(protocols/groth16/prove.py:133-248), sequential re-use of one key.
"""
from __future__ import annotations

import heapq
import importlib
import threading
import time
from collections import deque
from concurrent.futures import Future
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
    admitted: float = 0.0
    done: float = 0.0
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _HostPool:
    """Host workers for the G2 MSMs, shortest job first: a queued job with
    fewer bases runs before a larger one, so a small proof does not wait for
    the G2 MSMs of large proofs admitted before it."""

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
    [w | h | l] plus the per-batch a, b, c and scratch vectors (counted per
    request, an upper bound)."""
    return P.segment_bytes(pk) + 32 * 4 * pk.n


class ProveServer:
    def __init__(self, keys: Dict[object, ProvingKey], hbm_budget: int, max_batch: int = 16,
                 g2_workers: int = 2, cpu_threads: int = 16) -> None:
        self.keys = keys
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.pool = _HostPool(g2_workers)
        self.cpu_threads = cpu_threads
        self.stats = {"batches": 0, "msm_launches": 0, "max_batch_seen": 0}

    def close(self) -> None:
        self.pool.shutdown()

    def footprint(self, req: Request) -> int:
        return working_set_bytes(self.keys[req.key])

    def run(self, requests: List[Request]) -> List[Result]:
        """Feeds `requests` at their arrival times, serves them and returns
        the results in request order."""
        reqs = sorted(requests, key=lambda q: q.arrival)
        for q in reqs:
            if q.key not in self.keys:
                raise ValueError(f"request {q.id}: unknown circuit {q.key!r}")
            if self.footprint(q) > self.budget:
                raise ValueError(f"request {q.id} needs {self.footprint(q)} device bytes, above the HBM budget "
                                 f"{self.budget}")
        results: Dict[int, Result] = {}
        res_lock = threading.Lock()
        queue: deque = deque()
        cv = threading.Condition()
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0

        def feeder():
            for q in reqs:
                delay = q.arrival - now()
                if delay > 0:
                    time.sleep(delay)
                with cv:
                    queue.append(q)
                    cv.notify()

        def finish(q: Request, admitted: float, handoff: dict, ready: threading.Event):
            pk = self.keys[q.key]
            err = None
            proof = None
            try:
                b2 = P.g2_part(pk, q.witness, self.cpu_threads)
                ready.wait()
                if "error" in handoff:
                    err = handoff["error"]
                else:
                    proof = P.assemble(pk, handoff["g1"], b2, q.r, q.s)
            except Exception as e:          # recorded per request, the server keeps going
                err = f"{type(e).__name__}: {e}"
            with res_lock:
                results[q.id] = Result(q.id, proof, q.arrival, admitted, now(), err)

        ft = threading.Thread(target=feeder, daemon=True)
        ft.start()
        futures = []
        pending: List[tuple] = []
        try:
            self._schedule(reqs, queue, cv, now, finish, futures, pending)
        finally:
            # an exception in the scheduler must not leave pool tasks waiting
            for hand, ev in pending:
                if not ev.is_set():
                    hand.setdefault("error", "server aborted")
                    ev.set()
        for f in futures:
            f.result()
        ft.join()
        return [results[q.id] for q in requests]

    def _schedule(self, reqs, queue, cv, now, finish, futures, pending) -> None:
        admitted_count = 0
        while admitted_count < len(reqs):
            with cv:
                while not queue:
                    cv.wait()
                batch, used = [], 0
                while queue and len(batch) < self.max_batch:
                    fp = self.footprint(queue[0])
                    if batch and used + fp > self.budget:
                        break
                    batch.append(queue.popleft())
                    used += fp
            t_adm = now()
            admitted_count += len(batch)
            self.stats["batches"] += 1
            self.stats["max_batch_seen"] = max(self.stats["max_batch_seen"], len(batch))
            handoffs = {q.id: ({}, threading.Event()) for q in batch}
            pending.extend(handoffs.values())
            for q in batch:
                futures.append(self.pool.submit(self.keys[q.key].n_vars, finish, q, t_adm, *handoffs[q.id]))
            groups: Dict[object, List[Request]] = {}
            for q in batch:
                groups.setdefault(q.key, []).append(q)
            for key, qs in groups.items():
                try:
                    parts = P.device_batch(self.keys[key], [q.witness for q in qs])
                    self.stats["msm_launches"] += 1
                    for q, part in zip(qs, parts):
                        handoffs[q.id][0]["g1"] = part
                except ValueError:
                    # one bad witness fails its whole group launch: redo the group one by one
                    for q in qs:
                        try:
                            handoffs[q.id][0]["g1"] = P.device_batch(self.keys[key], [q.witness])[0]
                            self.stats["msm_launches"] += 1
                        except ValueError as e:
                            handoffs[q.id][0]["error"] = str(e)
                for q in qs:
                    handoffs[q.id][1].set()


def serve_plain(keys: Dict[object, ProvingKey], requests: List[Request], cpu_threads: int = 16) -> List[Result]:
    """Baseline: one proof at a time in arrival order, each with prove() (its
    G2 MSM on `cpu_threads` host threads, overlapped only with its own
    device work)."""
    t0 = time.perf_counter()
    out = {}
    for q in sorted(requests, key=lambda x: x.arrival):
        delay = q.arrival - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        start = time.perf_counter() - t0
        try:
            pf = P.prove(keys[q.key], q.witness, q.r, q.s, g2_threads=cpu_threads)
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
