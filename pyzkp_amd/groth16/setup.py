"""Groth16 trusted setup on the device (reference: protocols/groth16/setup.py:25-121).

Same key as the reference for the same trapdoor (tau, alpha, beta, gamma,
delta), computed without per-variable interpolation:
  * Lagrange values l_j = L_j(tau) = omega^j (tau^n - 1) / (n (tau - omega^j))
    on the device (powers, one batch of inversions);
  * u_i(tau) = sum_j A[j][i] l_j etc. as one transposed sparse product
    (the reference interpolates every column with an iNTT and evaluates it
    at tau: setup.py:63-73 — the same numbers);
  * G1 points [x] G1 by the device fixed-base kernel, G2 points [x] G2 on
    host threads.
The trapdoor is drawn from a seeded stream (the reference uses secrets,
setup.py:32-45), so keys and proofs are reproducible.
"""
from __future__ import annotations

import numpy as np

from .. import workload
from ..curve import G1_GENERATOR, fixed_base_table, g1_mul
from ..field import FR_GENERATOR, FR_MODULUS, fr_inv, fr_to_mont, omega_for_size, words_to_ints
from .. import g2 as G2
from .synthetic import CSR, SyntheticCircuit
from .types import ProvingKey, VerifyingKey

_TABLES: dict = {}


def _g1_table(ctx):
    t = _TABLES.get(id(ctx))
    if t is None:
        t = ctx.upload(fixed_base_table())
        _TABLES[id(ctx)] = t
    return t


def draw_trapdoor(seed: int):
    """(tau, alpha, beta, gamma, delta), non-zero, with tau^n != 1 checked by the caller."""
    v = words_to_ints(workload.fr_uniform(seed, 5))
    return [x if x else 1 for x in v]


def _upload_csr(ctx, m: CSR):
    rp = ctx.upload(np.ascontiguousarray(m.row_ptr, dtype="<u4").tobytes())
    cols = ctx.upload(np.ascontiguousarray(m.cols, dtype="<u4").tobytes() or b"\0" * 4)
    co = ctx.upload(np.ascontiguousarray(m.coeffs, dtype="<u8").tobytes() or b"\0" * 32)
    if len(m.cols):
        ctx.fr_to_mont(co.ptr, co.ptr, len(m.cols))
    return rp, cols, co


def setup(circ: SyntheticCircuit, ctx, seed: int = 0x5EED, window_bits: int = 0) -> ProvingKey:
    n, nv = circ.n, circ.n_vars
    omega = omega_for_size(n)
    tau, alpha, beta, gamma, delta = draw_trapdoor(seed)
    while (pow(tau, n, FR_MODULUS) - 1) % FR_MODULUS == 0:  # setup.py:35-39
        tau = (tau + 1) % FR_MODULUS
    t_tau = (pow(tau, n, FR_MODULUS) - 1) % FR_MODULUS
    inv_gamma, inv_delta = fr_inv(gamma), fr_inv(delta)

    # l_j = L_j(tau)
    om = ctx.alloc(32 * n)
    ctx.fr_powers(fr_to_mont(omega), fr_to_mont(1), n, om.ptr)
    tv = ctx.alloc(32 * n)
    ctx.fr_powers(fr_to_mont(1), fr_to_mont(tau), n, tv.ptr)          # all tau
    ctx.fr_sub(tv.ptr, om.ptr, tv.ptr, n)                              # tau - omega^j
    ctx.fr_inv(tv.ptr, tv.ptr, n)
    ctx.fr_mul(tv.ptr, om.ptr, tv.ptr, n)
    ctx.fr_scale(tv.ptr, fr_to_mont(t_tau * fr_inv(n) % FR_MODULUS), tv.ptr, n)
    lag = tv
    om.free()

    csr = {}
    polys = {}
    for name, mat in (("A", circ.A), ("B", circ.B), ("C", circ.C)):
        csr[name] = _upload_csr(ctx, mat)
        t = mat.transpose(nv)
        rp, cols, co = _upload_csr(ctx, t)
        out = ctx.alloc(32 * nv)
        ctx.r1cs_eval(rp.ptr, cols.ptr, co.ptr, lag.ptr, nv, out.ptr)
        polys[name] = out
        for b in (rp, cols, co):
            b.free()
    lag.free()
    u, v, w = polys["A"], polys["B"], polys["C"]
    # K_i = beta u_i + alpha v_i + w_i
    K = ctx.alloc(32 * nv)
    tmp = ctx.alloc(32 * nv)
    ctx.fr_scale(u.ptr, fr_to_mont(beta), K.ptr, nv)
    ctx.fr_scale(v.ptr, fr_to_mont(alpha), tmp.ptr, nv)
    ctx.fr_add(K.ptr, tmp.ptr, K.ptr, nv)
    ctx.fr_add(K.ptr, w.ptr, K.ptr, nv)

    aux = np.ascontiguousarray(circ.aux_ids, dtype="<u4")
    n_aux = len(aux)
    aux_dev = ctx.upload(aux.tobytes() or b"\0" * 4)
    n_h = n - 1
    total = 2 * nv + n_h + n_aux
    scal = ctx.alloc(32 * max(total, 1))
    ctx.d2d(scal.ptr, u.ptr, 32 * nv)
    ctx.d2d(scal.ptr + 32 * nv, v.ptr, 32 * nv)
    # h_query scalars: tau^k t(tau) / delta, k < n - 1   (setup.py:85-89)
    if n_h:
        ctx.fr_powers(fr_to_mont(tau), fr_to_mont(t_tau * inv_delta % FR_MODULUS), n_h, scal.ptr + 32 * 2 * nv)
    # l_query scalars: K_i / delta over aux ids
    if n_aux:
        ctx.fr_gather(K.ptr, aux_dev.ptr, n_aux, tmp.ptr)
        ctx.fr_scale(tmp.ptr, fr_to_mont(inv_delta), scal.ptr + 32 * (2 * nv + n_h), n_aux)
    pts = ctx.alloc(64 * max(total, 1))
    ctx.g1_fixed_base(_g1_table(ctx).ptr, scal.ptr, total, pts.ptr)
    ctx.sync()
    fixed = ctx.fixed_bases_create(pts.ptr, total, window_bits)
    csr_arrays = {name: (m.row_ptr, m.cols, m.coeffs) for name, m in (("A", circ.A), ("B", circ.B), ("C", circ.C))}

    # public inputs: ic_i = [K_i / gamma] G1   (setup.py:77-79, 91-93)
    ctx.fr_from_mont(K.ptr, tmp.ptr, nv)
    kvals = words_to_ints(np.frombuffer(ctx.download(tmp.ptr, 32 * nv), dtype="<u8").reshape(nv, 4))
    ic = [g1_mul(G1_GENERATOR, kvals[i] * inv_gamma % FR_MODULUS) for i in circ.public_ids]
    # b_query_g2 = [v_i(tau)] G2 on host threads
    ctx.fr_from_mont(v.ptr, tmp.ptr, nv)
    vw = np.frombuffer(ctx.download(tmp.ptr, 32 * nv), dtype="<u8").reshape(nv, 4).copy()
    b_g2 = G2.from_words(G2.g2_mul_many_words(G2.G2_GENERATOR, vw))
    for b in (u, v, w, K, tmp, scal):
        b.free()

    vk = VerifyingKey(alpha_g1=g1_mul(G1_GENERATOR, alpha), beta_g2=G2.g2_mul(G2.G2_GENERATOR, beta),
                      gamma_g2=G2.g2_mul(G2.G2_GENERATOR, gamma), delta_g2=G2.g2_mul(G2.G2_GENERATOR, delta),
                      ic=ic, public_names=["ONE"] + [f"v{i}" for i in circ.public_ids[1:]])
    pk = ProvingKey(vk=vk, beta_g1=g1_mul(G1_GENERATOR, beta), delta_g1=g1_mul(G1_GENERATOR, delta),
                    aux_ids=circ.aux_ids, trapdoor_tau=tau, trapdoor_alpha=alpha, trapdoor_beta=beta,
                    trapdoor_gamma=gamma, trapdoor_delta=delta, n=n, n_vars=nv, public_ids=list(circ.public_ids),
                    ctx=ctx, fixed=fixed, query_offsets=(0, nv, 2 * nv, 2 * nv + n_h), query_points=pts,
                    b_query_g2=b_g2, csr=csr, aux_dev=aux_dev, query_total=total)
    pk.csr_host = csr_arrays
    pk.hbm_bytes = key_hbm_bytes(pk)
    return pk


def key_hbm_bytes(pk: ProvingKey) -> int:
    W = (255 + pk.fixed.window_bits - 1) // pk.fixed.window_bits
    csr_bytes = sum(len(rp) * 4 + len(cols) * 36 for rp, cols, _ in pk.csr_host.values())
    return 64 * pk.query_total * (W + 1) + csr_bytes + 4 * len(pk.aux_ids)
