"""Served Groth16 on the device: every proof of a Poisson mix of circuit
sizes verifies, equals byte for byte the proof the plain one-at-a-time loop
makes for the same request (G2 MSMs on host workers or on the device
slots), and a witness that does not satisfy its R1CS
fails only its own request."""
import pytest

from pyzkp_amd.field import words_to_ints
from pyzkp_amd.groth16 import setup, synthetic_circuit, verify
from pyzkp_amd.groth16.prove import release_work
from pyzkp_amd.serve import (ProveServer, latency_stats, make_requests, poisson_arrivals, request_classes,
                             serve_plain, working_set_bytes)

pytestmark = pytest.mark.gpu


def test_served_equals_plain_and_verifies(zctx):
    classes = request_classes(12, 6, 10)
    circuits = {k: synthetic_circuit(k, 0x5300 + k) for k in set(classes)}
    keys = {k: setup(circuits[k], zctx, seed=0x5400 + k, window_bits=8) for k in circuits}
    reqs = make_requests(circuits, classes, poisson_arrivals(len(classes), 20000.0))
    bad = reqs[3].witness.copy()
    bad[0, 0] ^= 1
    reqs[3].witness = bad
    plain = serve_plain(keys, reqs, cpu_threads=4)
    # budget for about three requests per batch
    budget = 3 * max(working_set_bytes(pk) for pk in keys.values())
    srv = ProveServer(keys, hbm_budget=budget, max_batch=4, g2_workers=4, cpu_threads=4)
    served = srv.run(reqs)
    srv.close()
    assert srv.stats["max_batch_seen"] >= 2
    for q, a, b in zip(reqs, served, plain):
        if q.id == 3:
            assert a.proof is None and b.proof is None and "R1CS" in a.error
            continue
        assert a.error is None and a.proof is not None
        assert a.proof.to_bytes() == b.proof.to_bytes()
        pub = words_to_ints(q.witness[circuits[q.key].public_ids])
        assert verify(keys[q.key].vk, pub, a.proof)
    assert latency_stats(served)["count"] == len(reqs) - 1
    # the same requests with the G2 MSMs on the device slots
    srv = ProveServer(keys, hbm_budget=budget, max_batch=4, cpu_threads=4, g2_device=True)
    served_dev = srv.run(reqs)
    srv.close()
    for q, a, b in zip(reqs, served_dev, plain):
        if q.id == 3:
            assert a.proof is None and "R1CS" in a.error
            continue
        assert a.error is None and a.proof.to_bytes() == b.proof.to_bytes()
    with pytest.raises(ValueError):
        ProveServer(keys, hbm_budget=1024).run(reqs[:1])
    for pk in keys.values():
        release_work(pk)
        pk.free()
