"""Scheduler logic of the prove server on the CPU, with the device and G2
steps replaced by stand-ins and stand-in slot contexts: every request is
answered once, groups respect the HBM budget and the batch cap, groups of
different circuits run on different slots at the same time, requests that
queue for G2 workers are batched on the device, a failing witness fails only
its own request, and an over-budget request is refused before the run."""
import threading
import time

import numpy as np
import pytest

import pyzkp_amd.serve as S


class _Key:
    def __init__(self, n, nv):
        self.n, self.n_vars, self.aux_ids = n, nv, np.arange(nv - 2)


class _Ctx:
    def __init__(self, i):
        self.i = i

    def sync(self, stream=None):
        pass

    def close(self):
        pass


class _Work:
    def __init__(self, ctx, k):
        self.ctx, self.k, self.nbytes = ctx, k, 0

    def free(self):
        pass


@pytest.fixture
def fake(monkeypatch):
    calls = {"device": [], "lock": threading.Lock(), "active": 0, "max_active": 0, "keys_live": set(),
             "cross": False, "dev_s": 0.0, "device_g2": 0}

    def device_batch(pk, ws, work, g2=False):
        with calls["lock"]:
            calls["device"].append((pk.n, len(ws), work.ctx.i))
            calls["active"] += 1
            calls["max_active"] = max(calls["max_active"], calls["active"])
            if calls["keys_live"] and pk.n not in calls["keys_live"]:
                calls["cross"] = True           # two circuits on the device at once
            calls["keys_live"].add(pk.n)
        assert work.k >= len(ws)
        time.sleep(calls["dev_s"])
        with calls["lock"]:
            calls["active"] -= 1
            calls["keys_live"].discard(pk.n)
        for w in ws:
            if int(w[0, 0]) == 0xBAD:
                raise ValueError("witness does not satisfy R1CS (non-zero remainder)")
        if g2:
            calls["device_g2"] += len(ws)
            return [(("g1", int(w[0, 0])), ("g2", int(w[0, 0]))) for w in ws]
        return [("g1", int(w[0, 0])) for w in ws]

    def grow_work(pk, w, k, ctx=None):
        return w if w is not None and w.k >= k and w.ctx is ctx else _Work(ctx, max(k, w.k if w else 0))

    def g2_part(pk, w, threads):
        time.sleep(0.001 * (pk.n // 64))
        return ("g2", int(w[0, 0]))

    def assemble(pk, g1, g2, r, s):
        assert g1[1] == g2[1]
        return ("proof", g1[1], r, s)

    monkeypatch.setattr(S.P, "device_batch", device_batch)
    monkeypatch.setattr(S.P, "grow_work", grow_work)
    monkeypatch.setattr(S.P, "g2_part", g2_part)
    monkeypatch.setattr(S.P, "assemble", assemble)
    return calls


def _reqs(keys, classes, bad=(), gap=0.0005):
    out = []
    for i, k in enumerate(classes):
        w = np.zeros((keys[k].n_vars, 4), dtype=np.uint64)
        w[0, 0] = 0xBAD if i in bad else 1000 + i
        out.append(S.Request(id=i, key=k, witness=w, r=i, s=2 * i, arrival=gap * i))
    return out


def _server(keys, budget, **kw):
    return S.ProveServer(keys, hbm_budget=budget, slot_contexts=[_Ctx(0), _Ctx(1)], **kw)


def test_every_request_answered_within_budget(fake):
    keys = {6: _Key(64, 70), 8: _Key(256, 260)}
    classes = [6, 8, 6, 6, 8, 8, 6, 8, 6, 6, 8, 6, 6, 8, 6, 8]
    reqs = _reqs(keys, classes, bad={5})
    budget = 3 * S.working_set_bytes(keys[8])
    srv = _server(keys, budget, max_batch=4, g2_workers=3, cpu_threads=2)
    res = srv.run(reqs)
    srv.close()
    assert [r.id for r in res] == list(range(len(reqs)))
    for q, r in zip(reqs, res):
        if q.id == 5:
            assert r.proof is None and "R1CS" in r.error
        else:
            assert r.error is None and r.proof == ("proof", 1000 + q.id, q.id, 2 * q.id)
        assert r.done >= r.arrival - 1e-9
    per = {64: S.working_set_bytes(keys[6]), 256: S.working_set_bytes(keys[8])}
    assert all(k <= 4 for _, k, _ in fake["device"])
    assert srv.stats["max_batch_seen"] <= 4
    assert sum(k for _, k, _ in fake["device"]) >= len(reqs)
    assert max(per[n] * k for n, k, _ in fake["device"]) <= budget
    assert srv.stats["peak_inflight_bytes"] <= budget


def test_device_g2_mode_uses_no_host_workers(fake, monkeypatch):
    # g2_device: every G2 part comes from the slot's device batch, none from
    # the host pool; a burst still forms batches; a bad witness fails alone
    def no_host_g2(pk, w, threads):
        raise AssertionError("host G2 MSM called in device G2 mode")

    monkeypatch.setattr(S.P, "g2_part", no_host_g2)
    fake["dev_s"] = 0.01
    keys = {6: _Key(64, 70), 8: _Key(256, 260)}
    classes = [6, 8, 6, 6, 8, 8, 6, 8, 6, 6, 8, 6]
    reqs = _reqs(keys, classes, bad={4}, gap=0.0)
    srv = _server(keys, 1 << 40, max_batch=4, cpu_threads=2, g2_device=True)
    res = srv.run(reqs)
    srv.close()
    for q, r in zip(reqs, res):
        if q.id == 4:
            assert r.proof is None and "R1CS" in r.error
        else:
            assert r.error is None and r.proof == ("proof", 1000 + q.id, q.id, 2 * q.id)
    assert fake["device_g2"] >= len(reqs) - 1
    assert srv.stats["max_batch_seen"] >= 2 and srv.stats["max_batch_seen"] <= 4


def test_groups_of_different_circuits_overlap_and_queue_forms_batches(fake):
    fake["dev_s"] = 0.02
    keys = {6: _Key(64, 70), 8: _Key(1024, 1030)}
    # a burst: G2 work (one worker) is the bottleneck, so requests wait for it
    classes = [8, 6] * 10
    reqs = _reqs(keys, classes, gap=0.0)
    srv = _server(keys, 1 << 40, max_batch=8, g2_workers=1, cpu_threads=1)
    res = srv.run(reqs)
    srv.close()
    assert all(r.error is None for r in res)
    assert fake["cross"] and fake["max_active"] == 2          # both slots busy with two circuits
    assert srv.stats["max_concurrent_groups"] == 2
    assert srv.stats["max_batch_seen"] >= 2
    assert {c for _, _, c in fake["device"]} == {0, 1}


def test_light_load_runs_each_request_alone(fake):
    keys = {6: _Key(64, 70)}
    reqs = _reqs(keys, [6] * 5, gap=0.05)
    srv = _server(keys, 1 << 40, max_batch=8, g2_workers=2, cpu_threads=1)
    res = srv.run(reqs)
    srv.close()
    assert all(r.error is None for r in res)
    assert srv.stats["max_batch_seen"] == 1 and srv.stats["groups"] == 5


def test_bad_witness_shape_fails_only_itself(fake):
    keys = {6: _Key(64, 70)}
    reqs = _reqs(keys, [6, 6, 6])
    reqs[1].witness = reqs[1].witness[:-1]
    srv = _server(keys, 1 << 40)
    res = srv.run(reqs)
    srv.close()
    assert res[1].proof is None and "witness must have shape" in res[1].error
    assert res[0].error is None and res[2].error is None


def test_over_budget_refused(fake):
    keys = {6: _Key(64, 70)}
    srv = _server(keys, 16)
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


def test_host_g2_failure_fails_only_its_request(fake, monkeypatch):
    real = S.P.g2_part

    def g2_part(pk, w, threads):
        if int(w[0, 0]) == 1003:
            raise RuntimeError("host G2 failure")
        return real(pk, w, threads)

    monkeypatch.setattr(S.P, "g2_part", g2_part)
    keys = {6: _Key(64, 70)}
    reqs = _reqs(keys, [6] * 6)
    srv = _server(keys, 1 << 40, max_batch=2)
    res = srv.run(reqs)
    srv.close()
    assert res[3].proof is None and "host G2 failure" in res[3].error
    assert all(r.error is None for i, r in enumerate(res) if i != 3)
    assert srv.stats["groups"] >= 1
