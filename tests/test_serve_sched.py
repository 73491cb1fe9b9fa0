"""Scheduler logic of the prove server on the CPU, with the device and G2
steps replaced by stand-ins: every request is answered once, batches respect
the HBM budget and the batch cap, a failing witness fails only its own
request, and an over-budget request is refused before the run."""
import threading
import time

import numpy as np
import pytest

import pyzkp_amd.serve as S


class _Key:
    def __init__(self, n, nv):
        self.n, self.n_vars, self.aux_ids = n, nv, np.arange(nv - 2)


@pytest.fixture
def fake(monkeypatch):
    calls = {"device": [], "lock": threading.Lock()}

    def device_batch(pk, ws):
        calls["device"].append((pk.n, len(ws)))
        for w in ws:
            if int(w[0, 0]) == 0xBAD:
                raise ValueError("witness does not satisfy R1CS (non-zero remainder)")
        return [("g1", int(w[0, 0])) for w in ws]

    def g2_part(pk, w, threads):
        time.sleep(0.001 * (pk.n // 64))
        return ("g2", int(w[0, 0]))

    def assemble(pk, g1, g2, r, s):
        assert g1[1] == g2[1]
        return ("proof", g1[1], r, s)

    monkeypatch.setattr(S.P, "device_batch", device_batch)
    monkeypatch.setattr(S.P, "g2_part", g2_part)
    monkeypatch.setattr(S.P, "assemble", assemble)
    return calls


def _reqs(keys, classes, bad=()):
    out = []
    for i, k in enumerate(classes):
        w = np.zeros((keys[k].n_vars, 4), dtype=np.uint64)
        w[0, 0] = 0xBAD if i in bad else 1000 + i
        out.append(S.Request(id=i, key=k, witness=w, r=i, s=2 * i, arrival=0.0005 * i))
    return out


def test_every_request_answered_within_budget(fake):
    keys = {6: _Key(64, 70), 8: _Key(256, 260)}
    classes = [6, 8, 6, 6, 8, 8, 6, 8, 6, 6, 8, 6, 6, 8, 6, 8]
    reqs = _reqs(keys, classes, bad={5})
    budget = 3 * S.working_set_bytes(keys[8])
    srv = S.ProveServer(keys, hbm_budget=budget, max_batch=4, g2_workers=3, cpu_threads=2)
    res = srv.run(reqs)
    srv.close()
    assert [r.id for r in res] == list(range(len(reqs)))
    for q, r in zip(reqs, res):
        if q.id == 5:
            assert r.proof is None and "R1CS" in r.error
        else:
            assert r.error is None and r.proof == ("proof", 1000 + q.id, q.id, 2 * q.id)
        assert r.done >= r.admitted >= r.arrival - 1e-9
    # every device launch fits the budget and the cap
    per = {64: S.working_set_bytes(keys[6]), 256: S.working_set_bytes(keys[8])}
    assert all(k <= 4 for _, k in fake["device"])
    assert srv.stats["max_batch_seen"] <= 4
    assert sum(k for _, k in fake["device"]) >= len(reqs)
    assert max(per[n] * k for n, k in fake["device"]) <= budget


def test_over_budget_refused(fake):
    keys = {6: _Key(64, 70)}
    srv = S.ProveServer(keys, hbm_budget=16)
    with pytest.raises(ValueError):
        srv.run(_reqs(keys, [6]))
    srv.close()


def test_classes_and_arrivals_deterministic():
    a = S.request_classes(32, 10, 20)
    assert a == S.request_classes(32, 10, 20)
    assert all(10 <= k <= 20 for k in a)
    t = S.poisson_arrivals(32, 5.0)
    assert all(t[i] < t[i + 1] for i in range(31))
    assert 2.0 < t[-1] < 20.0
