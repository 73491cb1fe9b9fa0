#!/usr/bin/env python3
"""Benchmark of the prover hot path on one MI355X.

Default (``python bench.py``): one step = one BN254 G1 MSM over n = 2^22
device-resident points (uniform Fr scalars, bases B_i = [k_i] G1 built on
the device, seeds 0x3001 / 0x3002), result returned to the host.  The
window size c is autotuned in 13..18 before the warm-up.  The line carries
the VALU roofline of the dominant kernel (bucket accumulation) and a CPU
baseline (the C restatement of the reference's Pippenger, one thread on a
2^18-point prefix and all host threads on the full input, both checked
against the GPU result), and in `secondary` the two other legs of the
metric:

* ``ntt_2^24``: the Fr NTT at 2^24 in four variants (GB/s = 64 N / t), with
  its VALU and HBM rooflines and the C restatement of the reference NTT as
  CPU baseline;
* ``served_groth16``: the served Groth16 prover over 2^10..2^20-constraint
  requests: capacity, a sweep of Poisson arrival rates served by the
  continuous-batching server and by the plain one-proof-at-a-time loop, and
  the highest rate each sustains (at most its own capacity) with p99
  latency within the objective; every proof verified and equal to the
  plain loop's byte for byte.  The G2 MSMs run on the device by default
  (--g2-host: on host threads).

``--workload ntt`` / ``serve`` print those legs as the line;
``--workload batch`` the batched fixed-base MSMs + NTTs replayed as one
hipGraph; ``plonkq`` the PLONK quotient; ``copy`` the counter calibration.
Prints exactly one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# VALU peak of MI355X: 256 CUs x 64 lanes x 2.4 GHz peak engine clock for a
# full-rate VALU instruction (v_mad_u64_u32 is one), the same rate as the
# published 78.6 TFLOP/s FP64 vector peak (39.3 T FMA/s).  The library's
# probe reports the sustained rate next to it (about 38.6 T/s).
VALU_PEAK = 256 * 64 * 2.4e9
FQ_MUL_MADS = 136          # 8 x 32-bit-limb CIOS Montgomery product
MADD_FQ_MULS = 10          # XYZZ mixed addition 8M + 2S


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["msm", "ntt", "batch", "serve", "copy", "plonkq"], default="msm")
    ap.add_argument("--requests", type=int, default=128, help="serve: number of proof requests")
    ap.add_argument("--serve-log", type=str, default="10,20", help="serve: constraint range 2^lo..2^hi")
    ap.add_argument("--cpu-threads", type=int, default=16, help="serve: host threads for G2 MSMs")
    ap.add_argument("--hbm-budget-gb", type=float, default=16.0, help="serve: device budget for working sets")
    ap.add_argument("--g2-workers", type=int, default=2, help="serve: concurrent host G2 MSMs")
    ap.add_argument("--log-n", type=int, default=None, help="MSM: 22, NTT: 24")
    ap.add_argument("--window-bits", type=int, default=0, help="0: autotune in 13..18")
    ap.add_argument("--skewed", action="store_true", help="skewed scalars + repeated bases")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log", type=int, default=18)
    ap.add_argument("--no-ntt", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="MSM line: leave out the served-prover leg")
    ap.add_argument("--device-slots", type=int, default=2, help="serve: concurrent device slots")
    ap.add_argument("--g2-host", action="store_true",
                    help="serve: G2 MSMs on host threads (--cpu-threads, --g2-workers) instead of on the device")
    ap.add_argument("--loads", type=str, default="0.8,0.9,1.0",
                    help="serve: Poisson arrival rates as fractions of the plain loop's capacity")
    ap.add_argument("--server-loads", type=str, default="0.85,0.92,1.0",
                    help="serve: further arrival rates as fractions of the server's capacity")
    ap.add_argument("--slo-ms", type=float, default=400.0,
                    help="serve: p99 latency objective (about 2.5x a 2^20-constraint proof's plain service time)")
    return ap.parse_args()


def host_cpu_info() -> dict:
    """CPU model and the number of cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "usable_cores": usable, "os_cpu_count": os.cpu_count()}


def cpu_msm_baseline(ctx, bases_ptr: int, scalar_words, m: int):
    """The C restatement of the reference Pippenger (crypto/msm/pippenger.py),
    one thread, on m canonical points / scalars; returns (pts/s, seconds, result)."""
    import numpy as np
    from oracle import cref
    from oracle.msm import default_window_bits
    bc = ctx.alloc(64 * m)
    ctx.fq_from_mont(bases_ptr, bc.ptr, 2 * m)
    braw = np.frombuffer(ctx.download(bc.ptr, 64 * m), dtype="<u8")
    bc.free()
    sraw = np.ascontiguousarray(scalar_words[:m]).reshape(-1)
    t0 = time.perf_counter()
    res = cref.msm_pippenger_words(braw, sraw, m, default_window_bits(m), threads=1)
    dt = time.perf_counter() - t0
    return m / dt, dt, res


def signed_digit_pairs(words, c: int) -> int:
    """Number of nonzero signed c-bit digits of the scalars (canonical
    little-endian u64 words, one row per scalar): the (point, window) pairs
    the bucket accumulation adds."""
    import numpy as np

    w = np.ascontiguousarray(words, dtype=np.uint64)
    n = w.shape[0]
    W = (255 + c - 1) // c
    half, full = 1 << (c - 1), 1 << c
    carry = np.zeros(n, dtype=np.int64)
    count = 0
    for k in range(W):
        bit = c * k
        lo = np.zeros(n, dtype=np.uint64)
        for j in range(4):   # bits [bit, bit + c) spread over at most two words
            base = 64 * j
            if base + 64 <= bit or base >= bit + c:
                continue
            if base <= bit:
                lo |= w[:, j] >> np.uint64(bit - base)
            else:
                lo |= w[:, j] << np.uint64(base - bit)
        d = (lo & np.uint64(full - 1)).astype(np.int64) + carry
        carry = (d >= half).astype(np.int64)
        d -= carry * full
        count += int(np.count_nonzero(d))
    return count


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` for this workload, read from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a
    streaming copy kernel); None when that workload was not profiled."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)[workload][kernel]
    except (OSError, KeyError, ValueError):
        return None
    k = int(t.get("launches_per_transform", 1))   # an NTT transform is several pass launches
    return {"bytes": k * (t["read_bytes"] + t["write_bytes"]), "read_bytes": k * t["read_bytes"],
            "write_bytes": k * t["write_bytes"], "launches": k,
            "source": "profiles/hbm_traffic.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, calibrated)"}


def ntt_block(ctx, log_n, steps, warmup):
    """Times the four NTT variants at 2^log_n; returns a dict."""
    from pyzkp_amd import workload
    from pyzkp_amd.field import FR_GENERATOR, fr_to_mont

    n = 1 << log_n
    src = workload.device_fr_mont(ctx, workload.fr_uniform(0x2001 if log_n == 20 else 0x2002, n))
    dst = ctx.alloc(32 * n)
    shift = fr_to_mont(FR_GENERATOR)
    res = {}
    ctx.profile_enable(True)
    for name, inv, sh in (("fwd", False, None), ("inv", True, None), ("coset_fwd", False, shift),
                          ("coset_inv", True, shift)):
        for _ in range(warmup):
            ctx.ntt(src.ptr, n, dst.ptr, log_n, inv, shift_mont=sh)
        ctx.sync()
        ctx.profile_reset()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.ntt(src.ptr, n, dst.ptr, log_n, inv, shift_mont=sh)
        ctx.sync()
        wall = (time.perf_counter() - t0) / steps
        dev_ms, launches = ctx.profile_get("ntt")
        dev = dev_ms / max(launches, 1) / 1e3
        res[name] = {"ms": round(wall * 1e3, 4), "device_ms": round(dev * 1e3, 4),
                     "GBps": round(64 * n / wall / 1e9, 2)}
    ctx.profile_enable(False)
    src.free()
    dst.free()
    return res


def ntt_leg(ctx, log_n: int, steps: int, warmup: int, probe_mads: float, with_cpu: bool,
            cpu_threads: int = 16) -> dict:
    """Fr NTT at 2^log_n (config 2): the four variants timed, the VALU and
    HBM rooflines of the forward transform, and the CPU baseline (the C
    restatement of crypto/poly/ntt.py, forward at 2^min(22, log_n), one
    thread and all host threads)."""
    from pyzkp_amd import workload
    from pyzkp_amd.field import omega_for_size

    r = ntt_block(ctx, log_n, steps, warmup)
    n = 1 << log_n
    fwd = r["fwd"]
    dev_s = fwd["device_ms"] / 1e3
    mads = (n // 2) * log_n * FQ_MUL_MADS
    algo_bytes = 64 * n
    leg = {
        "metric": f"Fr NTT GB/s @2^{log_n}", "value": fwd["GBps"], "unit": "GB/s", "ms": fwd["ms"],
        "workload": f"Fr NTT forward n=2^{log_n}, uniform coefficients seed {'0x2001' if log_n == 20 else '0x2002'}",
        "variants": r,
        "roofline": {"bound": "valu", "achieved": round(mads / dev_s / 1e12, 3), "peak": round(VALU_PEAK / 1e12, 3),
                     "unit": "T lane-mads/s", "frac": round(mads / dev_s / VALU_PEAK, 4),
                     "peak_source": f"256 CUs x 64 lanes x 2.4 GHz; probe {probe_mads / 1e12:.2f} T/s sustained",
                     "work_per_launch": f"(n/2) log2 n = {(n // 2) * log_n} butterflies x 136 v_mad_u64_u32",
                     "kernel": "ntt_pass_kernel (all passes of one transform)",
                     "hbm": {"achieved": round(algo_bytes / dev_s / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                             "frac": round(algo_bytes / dev_s / 8e12, 4),
                             "algorithmic_bytes": algo_bytes},
                     "traffic": measured_traffic(f"ntt_2^{log_n}", "ntt_pass_kernel")},
        "cpu_baseline": None,
    }
    if with_cpu:
        from oracle import cref
        nc = 1 << min(22, log_n)
        lgc = nc.bit_length() - 1
        wds = workload.fr_uniform(0x2001, nc)
        t0 = time.perf_counter()
        cref.ntt_words(wds, nc, lgc, omega_for_size(nc), False, None, 1)
        c1 = time.perf_counter() - t0
        threads = max(1, min(cpu_threads, host_cpu_info()["usable_cores"] or 1))
        t0 = time.perf_counter()
        cref.ntt_words(wds, nc, lgc, omega_for_size(nc), False, None, threads)
        cm = time.perf_counter() - t0
        leg["cpu_baseline"] = {"value": round(64 * nc / c1 / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                               "sample": f"C restatement of crypto/poly/ntt.py (bit-reverse radix-2 DIT), forward "
                                         f"n=2^{lgc}, 1 thread, {c1:.2f} s",
                               "all_cores": {"value": round(64 * nc / cm / 1e9, 3), "unit": "GB/s",
                                             "cores": threads, "sample": f"the same, {threads} threads, {cm:.2f} s"}}
    return leg


def run_batch(args, ctx, peak_mads):
    """Config 4: 64 fixed-base MSMs, n_j = floor(2^U), U ~ U[10, 18] (seed
    0x4001), over one shared 2^18-point SRS, plus 64 forward NTTs with log2
    size uniform in 12..18 (seed 0x4002) grouped by size into batched jobs;
    one step = one hipGraph replay of the whole batch.  Compared with the
    same plan run eagerly and with the work issued one call at a time.  The
    roofline: the bucket accumulation's VALU rate (its device time from the
    eager plan's HIP events) and the whole step's."""
    from pyzkp_amd import workload
    from pyzkp_amd.curve import from_jacobian_mont
    from pyzkp_amd.field import words_to_ints
    from pyzkp_amd.tune import Tuner
    from oracle import msm as M

    wl = workload.batch_workload(ctx)
    sizes, offs, jobs = wl["sizes"], wl["offs"], wl["ntt_jobs"]
    nb, tot, ss = wl["srs_n"], sum(wl["sizes"]), wl["ss"]
    log("batch: SRS 2^18, 64 MSMs (%d points), 64 NTTs" % tot)
    ntt_elems = sum(j["batch"] << j["log_n"] for j in jobs)
    tuner = Tuner(ctx, use_cache=False)
    c, fb = tuner.fixed_window(wl["bases"].ptr, nb, wl["scalars"].ptr, offs, sizes, cands=range(11, 18))
    log(f"  fixed-base window c={c}: {tuner.log}")
    plans = {}
    acc_ms = None
    for mode in ("graph", "eager"):
        plan = ctx.batch_create(fb, wl["scalars"].ptr, offs, sizes, jobs, use_graph=(mode == "graph"))
        for _ in range(args.warmup):
            res = plan.run()
        if mode == "eager":
            ctx.profile_reset()
            ctx.profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = plan.run()
        plans[mode] = ((time.perf_counter() - t0) / args.steps, res)
        if mode == "eager":
            ctx.profile_enable(False)
            ms, cnt = ctx.profile_get("msm_accumulate")
            acc_ms = ms / max(args.steps, 1)
        plan.free()

    def plain():
        for o, m in zip(offs, sizes):
            ctx.msm_g1(wl["bases"].ptr, wl["scalars"].ptr + 32 * o, m, 0)
        for j in jobs:
            n = 1 << j["log_n"]
            for b in range(j["batch"]):
                ctx.ntt(j["in"] + 32 * n * b, n, j["out"] + 32 * n * b, j["log_n"], False)
        ctx.sync()
    plain()
    t0 = time.perf_counter()
    plain()
    t_plain = time.perf_counter() - t0

    log("closed-form check of 64 MSMs")
    kk, sv = words_to_ints(wl["ks"]), words_to_ints(ss)
    ok = True
    for j, (o, m) in enumerate(zip(offs, sizes)):
        p = from_jacobian_mont(*plans["graph"][1][j])
        ok &= (None if p[2] == 0 else (p[0], p[1])) == M.msm_closed_form(kk[:m], sv[o:o + m])
    step = plans["graph"][0]
    # algorithmic VALU work: every nonzero signed digit of every window is one
    # mixed addition (fixed-base tables: one bucket set for all windows)
    pairs = sum(signed_digit_pairs(ss[o:o + m], c) for o, m in zip(offs, sizes))
    msm_mads = pairs * MADD_FQ_MULS * FQ_MUL_MADS
    ntt_mads = sum(j["batch"] * ((1 << j["log_n"]) // 2) * j["log_n"] for j in jobs) * FQ_MUL_MADS
    acc_rate = msm_mads / (acc_ms / 1e3) if acc_ms else None
    line = {
        "metric": "batched fixed-base G1 MSM pts/s (64 MSMs, shared 2^18 SRS) + 64 NTTs, one hipGraph",
        "value": round(tot / step, 1), "unit": "pts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32x8 Montgomery Fq/Fr (254-bit integer)", "data": "synthetic",
        "config": {"workload": "64 MSMs n=floor(2^U), U~U[10,18] (seed 0x4001) over one 2^18 SRS (fixed-base "
                               "tables) + 64 forward NTTs 2^12..2^18 (seed 0x4002), grouped by size",
                   "msm_points": tot, "ntt_elements": ntt_elems, "ntt_jobs": len(jobs), "fixed_base_window": c,
                   "tune_ms": tuner.log, "graph_ms": round(plans["graph"][0] * 1e3, 4),
                   "eager_ms": round(plans["eager"][0] * 1e3, 4), "one_call_at_a_time_ms": round(t_plain * 1e3, 4),
                   "speedup_vs_one_at_a_time": round(t_plain / step, 3), "verified_closed_form": bool(ok)},
        "roofline": {"bound": "valu", "achieved": round(acc_rate / 1e12, 3) if acc_rate else None,
                     "peak": round(VALU_PEAK / 1e12, 3), "unit": "T lane-mads/s",
                     "frac": round(acc_rate / VALU_PEAK, 4) if acc_rate else None,
                     "kernel": "msm_accumulate_kernel (device time from the eager plan's HIP events)",
                     "accumulate_ms": round(acc_ms, 4) if acc_ms else None,
                     "work_per_launch": f"nonzero signed digits ({pairs}) * 10 Fq-mul * 136 = {msm_mads:.3e}",
                     "step_frac": round((msm_mads + ntt_mads) / step / VALU_PEAK, 4),
                     "step_work": f"MSM {msm_mads:.3e} + NTT {ntt_mads:.3e} v_mad_u64_u32 per replay",
                     "traffic": measured_traffic("batch", "msm_accumulate_kernel")},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        rate, dt, _ = cpu_msm_baseline(ctx, wl["bases"].ptr, ss, nb)
        line["cpu_baseline"] = {"value": round(rate, 1), "unit": "pts/s", "cores": 1, "kind": "port",
                                "sample": f"C restatement of crypto/msm/pippenger.py, one MSM over the whole "
                                          f"2^18-point SRS with the batch's first 2^18 scalars, 1 thread, "
                                          f"{dt:.2f} s"}
    print(json.dumps(line))
    fb.free()
    for b in wl["buffers"]:
        b.free()


def run_copy(args):
    """Counter calibration: a plain streaming kernel of known size (out = ~in
    over 2^28 int32 = 1 GiB read + 1 GiB written), so that the FETCH_SIZE /
    WRITE_SIZE counters of the hot kernels can be scaled to bytes."""
    import torch
    n = 1 << 28
    a = torch.randint(0, 1 << 30, (n,), dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        torch.bitwise_not(a, out=b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        torch.bitwise_not(a, out=b)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"metric": "streaming copy GB/s (counter calibration)", "value": round(2 * 4 * n / dt / 1e9, 1),
                      "unit": "GB/s", "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                      "ms_per_step": round(dt * 1e3, 4), "steps": args.steps, "warmup": args.warmup}))


def run_plonkq(args, ctx):
    """PLONK quotient (protocols/plonk/prove.py:300-424) at n = 2^log_n
    (default 2^20, extended coset m = 4n): one step = the 15 coset NTTs of
    the zero-padded coefficient vectors, the fused numerator / Z_H kernel,
    the batch inversion, the point-wise product and the coset iNTT, with
    the inputs resident in HBM (random coefficients, seed 0x6001).  The
    degree-check download is not part of the step."""
    import numpy as np
    from pyzkp_amd import workload
    from pyzkp_amd.field import FR_GENERATOR, FR_MODULUS, fr_to_mont, omega_for_size

    log_n = args.log_n or 20
    n, m = 1 << log_n, 1 << (log_n + 2)
    omega, omega_m = omega_for_size(n), omega_for_size(m)
    shift = FR_GENERATOR
    u = pow(omega_m, n, FR_MODULUS)
    assert pow(shift, n, FR_MODULUS) not in {1, u, u * u % FR_MODULUS, u * u * u % FR_MODULUS}
    log(f"plonk quotient n=2^{log_n}, m=2^{log_n + 2}")
    src = workload.device_fr_mont(ctx, workload.fr_uniform(0x6001, 14 * n))   # 13 inputs + L1
    ext = [ctx.alloc(32 * m) for _ in range(15)]
    num, zh, t = ctx.alloc(32 * m), ctx.alloc(32 * m), ctx.alloc(32 * m)
    sh_m, shw_m = fr_to_mont(shift), fr_to_mont(shift * omega % FR_MODULUS)
    consts = [fr_to_mont(v) for v in (omega_m, shift, pow(shift, n, FR_MODULUS), 3, 5, 7, 11, 13)]
    order = list(range(13)) + [3, 13]          # z(omega X) from z; L1 last

    def step():
        for k, j in enumerate(order):
            ctx.ntt(src.ptr + 32 * n * j, n, ext[k].ptr, log_n + 2, False,
                    shift_mont=shw_m if k == 13 else sh_m)
        ctx.plonk_quotient_evals([b.ptr for b in ext], n, m, *consts, num.ptr, zh.ptr)
        ctx.fr_inv(zh.ptr, zh.ptr, m)
        ctx.fr_mul(num.ptr, zh.ptr, t.ptr, m)
        ctx.ntt(t.ptr, m, t.ptr, log_n + 2, True, shift_mont=sh_m)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    dt = (time.perf_counter() - t0) / args.steps
    # the dominant kernel's device time, from HIP events on the NTT's stream,
    # in as many further steps outside the timed region
    ctx.profile_enable(True)
    ctx.profile_reset()
    for _ in range(args.steps):
        step()
    ctx.sync()
    ntt_ms, ntt_calls = ctx.profile_get("ntt")
    ctx.profile_enable(False)
    ntt_s = ntt_ms / 1e3 / args.steps
    ntt_mads = 16 * (m // 2) * (log_n + 2) * FQ_MUL_MADS
    # bytes: 16 NTTs of m (64 m each) + fused pass (17 vectors) + inversion (3 reads / 2 writes) + product (3)
    algo_bytes = 16 * 64 * m + 17 * 32 * m + 5 * 32 * m + 3 * 32 * m
    roofline = {
        "bound": "valu", "achieved": round(ntt_mads / ntt_s / 1e12, 3), "peak": round(VALU_PEAK / 1e12, 3),
        "unit": "T lane-mads/s", "frac": round(ntt_mads / ntt_s / VALU_PEAK, 4),
        "kernel": "ntt_pass_kernel (the 16 transforms of one step, HIP-event device time)",
        "work_per_launch": f"16 transforms x (m/2) log2 m = {16 * (m // 2) * (log_n + 2)} butterflies "
                           "x 136 v_mad_u64_u32",
        "ntt_ms_per_step": round(ntt_s * 1e3, 4), "ntt_calls_per_step": ntt_calls // max(args.steps, 1),
        "ntt_share_of_step": round(ntt_s / dt, 3),
        "hbm": {"achieved": round(16 * 64 * m / ntt_s / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(16 * 64 * m / ntt_s / 8e12, 4), "algorithmic_bytes": 16 * 64 * m},
        "traffic": None,
    }
    line = {
        "metric": "PLONK quotient (15 coset NTTs at 4n + fused numerator + batch inverse + coset iNTT) per second",
        "value": round(1.0 / dt, 3), "unit": "quotients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32 limbs: Fr Montgomery (254-bit integer)", "data": "synthetic",
        "config": {"workload": f"PLONK quotient n=2^{log_n} (m=2^{log_n + 2}), random coefficients seed 0x6001, "
                               "shift = 5", "algorithmic_GBps": round(algo_bytes / dt / 1e9, 1)},
        "roofline": roofline, "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        from oracle import plonk as OP
        from pyzkp_amd.field import words_to_ints
        nc = 1 << 11
        vals = words_to_ints(workload.fr_uniform(0x6001, 13 * nc))
        co = {k: vals[i * nc:(i + 1) * nc] for i, k in enumerate(OP.Q_INPUTS[:13])}
        t0 = time.perf_counter()
        OP.quotient_full(co, n=nc, omega=omega_for_size(nc), k1=7, k2=13, alpha=3, beta=5, gamma=7, shift=shift)
        cdt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(1.0 / cdt, 3), "unit": "quotients/s", "cores": 1, "kind": "port",
                                "sample": f"the oracle's restatement of protocols/plonk/prove.py:300-424 (Python), "
                                          f"n=2^11, 1 thread, {cdt:.2f} s: a smaller instance than the GPU's"}
    print(json.dumps(line))
    for b in ext + [num, zh, t, src]:
        b.free()


def oracle_key_from_device(pk, circ):
    """The oracle's proving-key dict for `pk`, read back from the device key
    (its G1 queries equal the oracle's setup point for point, which the GPU
    tests check), so the CPU prover can be timed without the oracle's slow
    setup."""
    from pyzkp_amd.curve import normalize
    from pyzkp_amd.g2 import normalize as normalize2, point_from_words
    vk = pk.vk
    return {"a_query": [normalize(p) for p in pk.g1_query("a")],
            "b_query_g1": [normalize(p) for p in pk.g1_query("b_g1")],
            "b_query_g2": [normalize2(point_from_words(pk.b_query_g2.words[i])) for i in range(circ.n_vars)],
            "h_query": [normalize(p) for p in pk.g1_query("h")],
            "l_query": [normalize(p) for p in pk.g1_query("l")],
            "alpha_g1": normalize(vk.alpha_g1), "beta_g1": normalize(pk.beta_g1),
            "beta_g2": normalize2(vk.beta_g2), "delta_g1": normalize(pk.delta_g1),
            "delta_g2": normalize2(vk.delta_g2), "aux_ids": [int(i) for i in circ.aux_ids], "n": circ.n}


def serve_leg(args, ctx, with_cpu: bool) -> dict:
    """Served Groth16 (config 5): Poisson arrivals (seed 0x5001) of requests
    whose circuit has floor(2^U) constraints, U ~ U[lo, hi] (seed 0x5002),
    padded to 2^k; one synthetic circuit and proving key per size class;
    witness seed 0x5100 + id, (r, s) from seed 0x5200 + id.

    * Capacity: all requests queued at t = 0, served by the plain
      one-at-a-time loop and by the continuous-batching server.
    * Rate sweep: the same request sequence arriving as a Poisson process at
      each of `loads` x the plain loop's capacity and `server_loads` x the
      server's, every rate served by both loops.  The sustained rate of a
      loop is the highest swept rate that is at most the loop's own capacity
      and whose p99 latency meets the objective (--slo-ms): above capacity a
      queue grows without bound, and a finite run meeting the objective there
      only ended before the queue did.
    Every server proof is checked with the pairing and must equal the plain
    loop's proof of the same request byte for byte (each request proved
    alone vs in the mix)."""
    from pyzkp_amd.field import words_to_ints
    from pyzkp_amd.groth16 import prove, setup, synthetic_circuit, synthetic_witness, verify
    from pyzkp_amd.serve import (ProveServer, latency_stats, make_requests, poisson_arrivals, request_classes,
                                 serve_plain, working_set_bytes)

    lo, hi = (int(x) for x in args.serve_log.split(","))
    R = args.requests
    g2d = not args.g2_host
    classes = request_classes(R, lo, hi)
    ks = sorted(set(classes))
    log(f"serve: {R} requests, classes {dict((k, classes.count(k)) for k in ks)}")
    t0 = time.perf_counter()
    circuits = {k: synthetic_circuit(k, 0x5300 + k) for k in ks}
    keys = {k: setup(circuits[k], ctx, seed=0x5400 + k) for k in ks}
    t_setup = time.perf_counter() - t0
    key_bytes = sum(pk.hbm_bytes for pk in keys.values())
    log(f"  keys ready in {t_setup:.1f} s ({key_bytes / 2**30:.2f} GiB resident)")
    svc = {}
    for k in ks:
        w = synthetic_witness(circuits[k], 0x5500 + k)
        prove(keys[k], w, 1, 2, g2_threads=args.cpu_threads, g2_device=g2d)
        t1 = time.perf_counter()
        prove(keys[k], w, 3, 4, g2_threads=args.cpu_threads, g2_device=g2d)
        svc[k] = time.perf_counter() - t1
    log(f"  plain service ms per class {dict((k, round(v * 1e3, 1)) for k, v in svc.items())}")
    budget = int(args.hbm_budget_gb * 2**30)
    checks = {"verified": 0, "identical_to_plain": 0, "proofs": 0}

    # one long-lived server, as in deployment: its slots build their NTT
    # tables and workspaces in a warm-up (two requests per class, a burst)
    srv = ProveServer(keys, hbm_budget=budget, max_batch=16, g2_workers=args.g2_workers,
                      cpu_threads=args.cpu_threads, device_slots=args.device_slots, g2_device=g2d)
    warm = make_requests(circuits, [k for k in ks for _ in range(2)], [0.0] * (2 * len(ks)), witness_seed=0x5600)

    def server_run(reqs):
        return srv.run(reqs), dict(srv.stats)

    def check(reqs, served, plain):
        for q, a, b in zip(reqs, served, plain):
            checks["proofs"] += 1
            if a.proof is None or b.proof is None:
                continue
            pub = words_to_ints(q.witness[circuits[q.key].public_ids])
            checks["verified"] += bool(verify(keys[q.key].vk, pub, a.proof))
            checks["identical_to_plain"] += a.proof.to_bytes() == b.proof.to_bytes()

    try:
        srv.run(warm)
        sat = make_requests(circuits, classes, [0.0] * R)
        log("  capacity: plain loop")
        plain_sat = serve_plain(keys, sat, cpu_threads=args.cpu_threads, g2_device=g2d)
        log("  capacity: server")
        srv_sat, sat_stats = server_run(sat)
        check(sat, srv_sat, plain_sat)
        cap_plain, cap_srv = latency_stats(plain_sat), latency_stats(srv_sat)
        base, base_srv = cap_plain["proofs_per_s"], cap_srv["proofs_per_s"]
        rates = sorted([(f * base, f"{f:g} x plain capacity") for f in (float(x) for x in args.loads.split(","))]
                       + [(f * base_srv, f"{f:g} x server capacity")
                          for f in (float(x) for x in args.server_loads.split(","))])
        sweep = []
        for rate, what in rates:
            reqs = make_requests(circuits, classes, poisson_arrivals(R, rate))
            log(f"  Poisson {rate:.2f}/s ({what}): plain, server")
            pl = serve_plain(keys, reqs, cpu_threads=args.cpu_threads, g2_device=g2d)
            sv, st = server_run(reqs)
            check(reqs, sv, pl)
            sweep.append({"load": round(rate / base, 4), "rate_of": what, "rate_per_s": round(rate, 3),
                          "plain": latency_stats(pl), "server": latency_stats(sv), "server_stats": st})
    finally:
        srv.close()
    slo = args.slo_ms

    def sustained(loop, cap=None):
        ok = [e["rate_per_s"] for e in sweep if e[loop].get("count") == R and e[loop]["p99_ms"] <= slo
              and (cap is None or e["rate_per_s"] <= cap * (1 + 1e-6))]
        return max(ok) if ok else 0.0

    sus_p, sus_s = sustained("plain", base), sustained("server", base_srv)
    at = [e for e in sweep if e["rate_per_s"] == sus_p]
    leg = {
        "metric": f"served Groth16: highest Poisson arrival rate, at most the loop's capacity, with p99 <= "
                  f"{slo:.0f} ms (server vs plain one-at-a-time loop), capacity, p50/p99",
        "value": sus_s, "unit": "proofs/s", "plain_value": sus_p,
        "rate_gain_at_equal_p99": round(sus_s / sus_p, 3) if sus_p else None,
        "finite_run_rate_within_slo": {"plain": sustained("plain"), "server": sustained("server")},
        "slo_p99_ms": slo,
        "p99_at_plain_sustained_rate": ({"plain": at[0]["plain"]["p99_ms"], "server": at[0]["server"]["p99_ms"]}
                                        if at else None),
        "workload": f"{R} requests, synthetic x^5 S-box R1CS 2^{lo}..2^{hi} constraints (log-uniform), "
                    "Poisson arrivals (seed 0x5001)",
        "classes": {str(k): classes.count(k) for k in ks},
        "service_ms_per_class": {str(k): round(v * 1e3, 2) for k, v in svc.items()},
        "capacity_server": cap_srv, "capacity_plain": cap_plain,
        "capacity_speedup": round(cap_srv["proofs_per_s"] / cap_plain["proofs_per_s"], 3),
        "server_stats_capacity": sat_stats, "sweep": sweep,
        "max_batch_seen_poisson": max(e["server_stats"]["max_batch_seen"] for e in sweep),
        "verified": checks["verified"], "identical_to_plain": checks["identical_to_plain"],
        "proofs_checked": checks["proofs"],
        "key_setup_s": round(t_setup, 2), "keys_hbm_gib": round(key_bytes / 2**30, 3),
        "working_set_max_mib": round(max(working_set_bytes(pk) for pk in keys.values()) / 2**20, 2),
        "cpu_threads": args.cpu_threads, "g2_workers": args.g2_workers, "device_slots": args.device_slots,
        "hbm_budget_gb": args.hbm_budget_gb, "g2_msm_on": "device" if g2d else "host threads",
        "cpu_baseline": None,
    }
    if with_cpu:
        from oracle import groth16 as OG
        k0 = ks[0]
        circ = circuits[k0]
        opk = oracle_key_from_device(keys[k0], circ)
        w = words_to_ints(synthetic_witness(circ, 0x5100))
        A, B, C = circ.A.row_dicts(), circ.B.row_dicts(), circ.C.row_dicts()
        log(f"  CPU baseline: the oracle's Groth16 prover, class 2^{k0}")
        t0 = time.perf_counter()
        OG.prove(opk, A, B, C, w, 3, 4)
        cdt = time.perf_counter() - t0
        leg["cpu_baseline"] = {"value": round(1.0 / cdt, 4), "unit": "proofs/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle's Groth16 prover (Python restatement of "
                                         f"protocols/groth16/prove.py), one proof of the smallest class 2^{k0}, "
                                         f"1 thread, {cdt:.2f} s (its key read back from the device key)"}
    for pk in keys.values():
        pk.free()
    return leg


def run_serve(args, ctx):
    leg = serve_leg(args, ctx, not args.no_cpu_baseline)
    line = {
        "metric": leg["metric"], "value": leg["value"], "unit": "proofs/s", "n_gpus": 1, "steps": args.requests,
        "warmup": 2 * len(leg["classes"]), "ms_per_step": round(1e3 / leg["value"], 3) if leg["value"] else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 limbs: Fr/Fq Montgomery (254-bit integer)", "data": "synthetic",
        "config": {k: v for k, v in leg.items() if k not in ("metric", "value", "cpu_baseline")},
        "roofline": None, "cpu_baseline": leg["cpu_baseline"],
    }
    print(json.dumps(line))
    return leg["verified"] == leg["proofs_checked"] and leg["identical_to_plain"] == leg["proofs_checked"]


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the MSM / NTT path here is single-device "
              "(multi-GPU sharding is a design note, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from pyzkp_amd import get_context, workload
    from pyzkp_amd.curve import G1_GENERATOR, from_jacobian_mont, g1_eq, g1_mul, normalize
    from pyzkp_amd.field import FR_MODULUS, words_to_ints

    if args.workload == "copy":
        run_copy(args)
        return

    ctx = get_context(0)
    log("probing v_mad_u64_u32 issue rate")
    probe_mads = ctx.probe_mad_rate()
    peak_mads = VALU_PEAK
    log(f"  {probe_mads / 1e12:.2f} T lane-mads/s sustained; peak {peak_mads / 1e12:.2f}")

    if args.workload == "batch":
        run_batch(args, ctx, peak_mads)
        return

    if args.workload == "plonkq":
        run_plonkq(args, ctx)
        return

    if args.workload == "serve":
        if not run_serve(args, ctx):
            sys.exit(1)
        return

    if args.workload == "ntt":
        log_n = args.log_n or 24
        leg = ntt_leg(ctx, log_n, args.steps, args.warmup, probe_mads, not args.no_cpu_baseline, args.cpu_threads)
        line = {
            "metric": leg["metric"], "value": leg["value"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": leg["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32 limbs: Fr Montgomery (254-bit integer)", "data": "synthetic",
            "config": {"workload": leg["workload"], "variants": leg["variants"]},
            "roofline": leg["roofline"], "cpu_baseline": leg["cpu_baseline"],
        }
        print(json.dumps(line))
        return

    log_n = args.log_n or 22
    n = 1 << log_n
    log(f"generating MSM inputs n=2^{log_n}{' (skewed)' if args.skewed else ''}")
    # the 2^12 single-MSM case uses its own seeds (0x1001 scalars, 0x1002 bases)
    seed_s, seed_k = (0x1001, 0x1002) if log_n == 12 else (0x3001, 0x3002)
    ks = workload.fr_uniform(seed_k, n)
    if args.skewed:
        ss = workload.scalars_skewed(0x3003, n)
        ks = workload.repeat_some(0x3004, ks, 0.05)
    else:
        ss = workload.fr_uniform(seed_s, n)
    bases = workload.device_bases(ctx, ks)
    scal = workload.device_fr_mont(ctx, ss)
    # the bases are a resident set (an SRS): converted once, like the MSM_G1
    # operator's per-points cache does
    mb = ctx.msm_bases_create(bases.ptr, n)
    msm = lambda m, c: ctx.msm_g1_prepared(mb, scal.ptr, m, c)

    # autotune c in 13..18 (one timed call each after one warm call)
    cands = [args.window_bits] if args.window_bits else list(range(13, 19))
    tune = {}
    # bucket reduction: digit split ("split"), or chunked running sums with
    # 0 (the library's rule), 8 or 16 buckets per thread
    reducers = (("split", 1, 0), ("chunk0", 0, 0), ("chunk8", 0, 8), ("chunk16", 0, 16))
    for c in cands:
        for name, split, k_red in reducers:
            ctx.set_tuning("msm_reduce_split", split)
            ctx.set_tuning("msm_reduce_chunk", k_red)
            msm(n, c)
            t0 = time.perf_counter()
            msm(n, c)
            msm(n, c)
            tune[(c, name)] = (time.perf_counter() - t0) / 2
            log(f"  c={c} reduce {name}: {tune[(c, name)] * 1e3:.2f} ms")
    c, red = min(tune, key=tune.get)
    _, split, k_red = next(r for r in reducers if r[0] == red)
    ctx.set_tuning("msm_reduce_split", split)
    ctx.set_tuning("msm_reduce_chunk", k_red)

    for _ in range(args.warmup):
        res = msm(n, c)
    ctx.sync()
    ctx.profile_reset()
    ctx.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = msm(n, c)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.profile_enable(False)
    phases = {}
    launches = {}
    for ph in ("msm_digits", "msm_sort", "msm_accumulate", "msm_fixup", "msm_reduce"):
        ms, cnt = ctx.profile_get(ph)
        # device time per MSM, summed over its window groups (groups on two
        # streams overlap, so the phases can add up to more than the step)
        phases[ph] = round(ms / args.steps, 4)
        launches[ph] = cnt // max(args.steps, 1)
    ms_per_step = wall / args.steps * 1e3
    value = n * args.steps / wall

    # full-size check: MSM = [sum s_i k_i mod r] G1
    log("closed-form check")
    kk = words_to_ints(ks)
    sv = words_to_ints(ss)
    t = sum(a * b for a, b in zip(kk, sv)) % FR_MODULUS
    verified = g1_eq(from_jacobian_mont(*res), g1_mul(G1_GENERATOR, t))

    W = (255 + c - 1) // c
    pairs = signed_digit_pairs(ss, c)
    algo_mads = pairs * MADD_FQ_MULS * FQ_MUL_MADS
    acc_s = phases["msm_accumulate"] / 1e3
    achieved = algo_mads / acc_s
    roofline = {"bound": "valu", "achieved": round(achieved / 1e12, 3), "peak": round(peak_mads / 1e12, 3),
                "unit": "T lane-mads/s", "frac": round(achieved / peak_mads, 4),
                "peak_source": f"256 CUs x 64 lanes x 2.4 GHz; probe {probe_mads / 1e12:.2f} T/s sustained",
                "traffic": None if args.skewed else measured_traffic(f"msm_2^{log_n}_c{c}", "msm_accumulate_kernel"),
                "kernel": "msm_accumulate_kernel",
                "launches_per_step": launches["msm_accumulate"],
                "work_per_launch": f"nonzero signed digits ({pairs} of n*W = {n * W}) * 10 Fq-mul * 136 "
                                   f"= {algo_mads:.3e} v_mad_u64_u32"}

    secondary = {}
    if not args.no_ntt:
        log("NTT 2^24")
        secondary["ntt_2^24"] = ntt_leg(ctx, 24, max(5, args.steps), 2, probe_mads, not args.no_cpu_baseline,
                                        args.cpu_threads)
    if not args.no_serve:
        secondary["served_groth16"] = serve_leg(args, ctx, not args.no_cpu_baseline)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import cref  # CPU baseline only (C restatement of the reference Pippenger)
        from oracle.msm import default_window_bits

        m = 1 << min(args.cpu_sample_log, log_n)
        log(f"CPU baseline: C Pippenger, n=2^{m.bit_length() - 1}, 1 thread")
        # canonical affine bases / scalars of the prefix, converted on the device
        bc = ctx.alloc(64 * m)
        ctx.fq_from_mont(bases.ptr, bc.ptr, 2 * m)
        braw = np.frombuffer(ctx.download(bc.ptr, 64 * m), dtype="<u8")
        bc.free()
        sraw = np.ascontiguousarray(ss[:m]).reshape(-1)
        wcpu = default_window_bits(m)
        t0 = time.perf_counter()
        cres = cref.msm_pippenger_words(braw, sraw, m, wcpu, threads=1)
        cpu_s = time.perf_counter() - t0
        gres = from_jacobian_mont(*msm(m, c))
        gaff = None if gres[2] == 0 else (gres[0], gres[1])
        cpu = {"value": round(m / cpu_s, 1), "unit": "pts/s", "cores": 1, "kind": "port",
               "sample": f"C restatement of crypto/msm/pippenger.py (unsigned {wcpu}-bit windows), "
                         f"first 2^{m.bit_length() - 1} points/scalars of the same inputs, 1 thread, "
                         f"{cpu_s:.2f} s",
               "gpu_match": gaff == cres}
        # the same restatement on the host's cores (OpenMP over windows), the full size
        threads = max(1, min(args.cpu_threads, host_cpu_info()["usable_cores"] or 1))
        log(f"CPU baseline: C Pippenger, n=2^{log_n}, {threads} threads")
        bc = ctx.alloc(64 * n)
        ctx.fq_from_mont(bases.ptr, bc.ptr, 2 * n)
        bfull = np.frombuffer(ctx.download(bc.ptr, 64 * n), dtype="<u8")
        bc.free()
        t0 = time.perf_counter()
        cfull = cref.msm_pippenger_words(bfull, np.ascontiguousarray(ss).reshape(-1), n, default_window_bits(n),
                                         threads=threads)
        mt_s = time.perf_counter() - t0
        full_aff = normalize(from_jacobian_mont(*res))
        cpu["all_cores"] = {"value": round(n / mt_s, 1), "unit": "pts/s", "cores": threads,
                            "sample": f"the same restatement, full 2^{log_n} inputs, {threads} threads, "
                                      f"{mt_s:.2f} s", "gpu_match": cfull == full_aff}
        cpu["host"] = host_cpu_info()

    line = {
        "metric": "G1 MSM pts/s @2^22; NTT GB/s @2^24; served proofs/s + p99 vs plain loop",
        "value": round(value, 1), "unit": "pts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32x8 Montgomery Fq/Fr (254-bit integer)", "data": "synthetic",
        "config": {"workload": f"BN254 G1 MSM n=2^{log_n} {'skewed' if args.skewed else 'uniform'} scalars, "
                               "bases [k_i]G1 affine (device-resident, prepared once as a resident set)",
                   "n": n, "window_bits": c, "reduce": red,
                   "autotune_ms": {f"c{k[0]}_{k[1]}": round(v * 1e3, 3) for k, v in tune.items()},
                   "phases_ms": phases, "verified_closed_form": bool(verified)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line))
    mb.free()
    bases.free()
    scal.free()


if __name__ == "__main__":
    main()
