"""The batched and served workloads at their full shapes.

Batched: 64 fixed-base MSMs, n_j = floor(2^U) with U ~ U[10, 18], over one
shared 2^18-point SRS, plus 64 forward NTTs of sizes 2^12..2^18 grouped by
size, recorded once into a hipGraph and replayed twice with eager work in
between; every MSM against the closed form [sum s_i k_i] G, every NTT of at
most 2^16 points against the C oracle in full, the 2^17 and 2^18 ones by
Horner rows and a device round trip.

Served: 48 Groth16 requests of 2^10..2^20 constraints (log-uniform), Poisson
arrivals, through ProveServer and through the plain one-at-a-time loop:
every proof verifies with the pairing, the served proof equals the plain
loop's byte for byte (the same request proved alone), and one request of
the smallest class equals the oracle's proof computed in the exponent from
the trapdoor."""
import numpy as np
import pytest

from oracle import cref
from oracle import groth16 as OG
from oracle import msm as M
from oracle import ntt as N
from oracle.field import R_MOD
from pyzkp_amd import workload
from pyzkp_amd.curve import from_jacobian_mont, normalize
from pyzkp_amd.field import words_to_ints
from pyzkp_amd.g2 import normalize as normalize2

pytestmark = pytest.mark.gpu


def _canon(ctx, ptr, n):
    tmp = ctx.alloc(32 * n)
    ctx.fr_from_mont(ptr, tmp.ptr, n)
    ctx.sync()
    w = np.frombuffer(ctx.download(tmp.ptr, 32 * n), dtype="<u8").reshape(n, 4).copy()
    tmp.free()
    return w


def test_batched_workload_full_shape(zctx):
    wl = workload.batch_workload(zctx)
    sizes, offs = wl["sizes"], wl["offs"]
    print(f"batched: {len(sizes)} MSMs, {sum(sizes)} points; NTT sizes {sorted(set(wl['ntt_logs']))}", flush=True)
    fb = zctx.fixed_bases_create(wl["bases"].ptr, wl["srs_n"], 15)
    plan = zctx.batch_create(fb, wl["scalars"].ptr, offs, sizes, wl["ntt_jobs"], use_graph=True)
    kk, sv = words_to_ints(wl["ks"]), words_to_ints(wl["ss"])
    expect = [M.msm_closed_form(kk[:m], sv[o:o + m]) for o, m in zip(offs, sizes)]
    threads = cref.max_threads()
    for rep in range(2):
        if rep:
            tmp = zctx.alloc(32 * (1 << 16))
            zctx.ntt(wl["scalars"].ptr, 1 << 16, tmp.ptr, 16, False)   # unrelated eager work
            zctx.sync()
            tmp.free()
        res = plan.run()
        for j in range(len(sizes)):
            p = from_jacobian_mont(*res[j])
            assert (None if p[2] == 0 else (p[0], p[1])) == expect[j], (rep, j)
        for job in wl["ntt_jobs"]:
            lg, cnt = job["log_n"], job["batch"]
            n = 1 << lg
            a = wl["ntt_inputs"][lg]
            out = _canon(zctx, job["out"], cnt * n)
            w = N.omega_for_size(n)
            for b in range(cnt):
                got = out[b * n:(b + 1) * n]
                coeffs = a[b * n:(b + 1) * n]
                if lg <= 16:
                    ref = cref.ntt_words(coeffs.reshape(-1), n, lg, w, False, None, threads)
                    assert np.array_equal(got.reshape(-1), ref), (rep, lg, b)
                    continue
                gi = words_to_ints(got)
                for i in (0, 1, n // 2 + 3, n - 1):
                    assert gi[i] == cref.poly_eval_words(coeffs, n, pow(w, i, R_MOD)), (rep, lg, b, i)
            if lg > 16:
                back = zctx.alloc(32 * cnt * n)
                zctx.ntt_batch(job["out"], back.ptr, lg, cnt, True)
                assert np.array_equal(_canon(zctx, back.ptr, cnt * n), a), (rep, lg)
                back.free()
    plan.free()
    fb.free()
    for b in wl["buffers"]:
        b.free()


def test_served_groth16_full_shape(zctx):
    from pyzkp_amd.groth16 import setup, synthetic_circuit, verify
    from pyzkp_amd.serve import (ProveServer, latency_stats, make_requests, poisson_arrivals, request_classes,
                                 serve_plain)
    classes = request_classes(48, 10, 20)
    ks = sorted(set(classes))
    print(f"served: 48 requests over classes {ks}", flush=True)
    circuits = {k: synthetic_circuit(k, 0x5300 + k) for k in ks}
    keys = {k: setup(circuits[k], zctx, seed=0x5400 + k) for k in ks}
    reqs = make_requests(circuits, classes, poisson_arrivals(48, 12.0))
    plain = serve_plain(keys, reqs, cpu_threads=16)
    srv = ProveServer(keys, hbm_budget=16 << 30, max_batch=16, g2_workers=2, cpu_threads=16)
    try:
        served = srv.run(reqs)
    finally:
        srv.close()
    print("served", latency_stats(served), srv.stats, "plain", latency_stats(plain), flush=True)
    for q, a, b in zip(reqs, served, plain):
        assert a.error is None and b.error is None, (q.id, a.error, b.error)
        assert a.proof.to_bytes() == b.proof.to_bytes(), q.id
        pub = words_to_ints(q.witness[circuits[q.key].public_ids])
        assert verify(keys[q.key].vk, pub, a.proof), q.id
    # the smallest class against the oracle, in the exponent
    q = next(x for x in reqs if x.key == ks[0])
    circ, pk = circuits[q.key], keys[q.key]
    td = (pk.trapdoor_tau, pk.trapdoor_alpha, pk.trapdoor_beta, pk.trapdoor_gamma, pk.trapdoor_delta)
    ek = OG.exponent_key(circ.A.row_dicts(), circ.B.row_dicts(), circ.C.row_dicts(), circ.n, circ.n_vars,
                         circ.public_ids, td)
    pf = served[q.id].proof
    got = {"a": normalize(pf.a), "b": normalize2(pf.b), "c": normalize(pf.c)}
    assert OG.check_in_exponent(ek, words_to_ints(q.witness), q.r, q.s, got, circ.n)
    for pk in keys.values():
        pk.free()
