"""Groth16 prover on the device (reference: protocols/groth16/prove.py:30-128,
prove_batch :133-248, QAP: protocols/groth16/qap.py:53-101).

Per proof, on the device: the witness (canonical words) is converted to
Montgomery form; A w, B w, C w are evaluated over the padded domain; a
witness that does not satisfy the R1CS raises ValueError (the reference's
non-zero remainder); H comes from 3 iNTTs, 3 coset NTTs (shift g = 5),
(a b - c) / (g^n - 1) point-wise and one coset iNTT — the same quotient
coefficients as the reference's poly_mul_ntt + div_by_xn_minus_1, since
Z_H(g w^i) = g^n - 1 on the coset.  The four G1 MSMs (a_query . w,
b_query_g1 . w, h_query . h[:n-1], l_query . w[aux]) run as one fixed-base
batch over the key's query set.  The G2 MSM (b_query_g2 . w) runs either
on host threads concurrently with the device work, or (g2_device) on the
device after the G1 batch, reading w where the G1 MSMs read it.  r and s are
arguments so proofs are reproducible (the reference draws them with fr_rand,
prove.py:84-85).
"""
from __future__ import annotations

import threading
from typing import List, Sequence, Tuple

import numpy as np

from .. import g2 as G2
from ..curve import from_jacobian_mont, g1_add, g1_mul_native as g1_mul
from ..field import FR_GENERATOR, FR_MODULUS, fr_inv, fr_to_mont
from .types import Proof, ProvingKey


class _Work:
    """Device buffers for k in-flight proofs of one key on one context.  A
    workspace records the key shape and the context it was sized for, and
    its owner (the key itself, or a serving slot) frees it."""

    def __init__(self, pk: ProvingKey, k: int, ctx=None) -> None:
        ctx = ctx if ctx is not None else pk.ctx
        n, nv, na = pk.n, pk.n_vars, len(pk.aux_ids)
        self.ctx = ctx
        self.n = n
        self.seg = nv + n + na              # scalars of one proof: [w | h | l]
        self.k = k
        self.S = ctx.alloc(32 * self.seg * k)
        self.abc = ctx.alloc(32 * 3 * n)
        self.tmp = ctx.alloc(32 * n)

    @property
    def nbytes(self) -> int:
        return 32 * (self.seg * self.k + 4 * self.n)

    def fits(self, pk: ProvingKey, k: int, ctx=None) -> bool:
        ctx = ctx if ctx is not None else pk.ctx
        return (self.S.ptr != 0 and self.ctx is ctx and self.n == pk.n and self.k >= k
                and self.seg == pk.n_vars + pk.n + len(pk.aux_ids))

    def free(self) -> None:
        for b in (self.S, self.abc, self.tmp):
            b.free()


def grow_work(pk: ProvingKey, w, k: int, ctx=None) -> _Work:
    """`w` if it holds k proofs of this key on ctx, else a fresh workspace
    (the old one freed once ctx is idle)."""
    if w is not None and w.fits(pk, k, ctx):
        return w
    kk = max(k, w.k if w is not None else 0)
    if w is not None:
        w.ctx.sync()
        w.free()
    return _Work(pk, kk, ctx)


def _work(pk: ProvingKey, k: int) -> _Work:
    """The key's own workspace with room for at least k proofs (grown, never
    shrunk)."""
    if pk.work is None or not pk.work.fits(pk, k):
        old, pk.work = pk.work, None
        pk.work = grow_work(pk, old, k)
    return pk.work


def segment_bytes(pk: ProvingKey) -> int:
    """Device bytes one in-flight proof of this key occupies ([w | h | l])."""
    return 32 * (pk.n_vars + pk.n + len(pk.aux_ids))


def release_work(pk: ProvingKey) -> None:
    w, pk.work = pk.work, None
    if w is not None:
        w.ctx.sync()
        w.free()


def _device_part(pk: ProvingKey, wk: _Work, slot: int, w_words: np.ndarray) -> None:
    """Fills scalars segment `slot` with [w | h | w[aux]] (Montgomery), on
    the workspace's context (its stream and arenas)."""
    ctx = wk.ctx
    n, nv = pk.n, pk.n_vars
    check_witness(pk, w_words)
    S = wk.S.ptr + 32 * wk.seg * slot
    wbuf = np.ascontiguousarray(w_words, dtype="<u8")   # alive until the synchronising check below
    ctx.h2d(S, wbuf.ctypes.data, 32 * nv)
    ctx.fr_to_mont(S, S, nv)
    a, b, c = wk.abc.ptr, wk.abc.ptr + 32 * n, wk.abc.ptr + 64 * n
    for name, out in (("A", a), ("B", b), ("C", c)):
        rp, cols, co = pk.csr[name]
        ctx.r1cs_eval(rp.ptr, cols.ptr, co.ptr, S, n, out)
    # R1CS satisfied on the domain <=> a b - c == 0 everywhere (qap.py:98-99)
    ctx.fr_mul(a, b, wk.tmp.ptr, n)
    ctx.fr_sub(wk.tmp.ptr, c, wk.tmp.ptr, n)
    bad = ctx.fr_count_nonzero(wk.tmp.ptr, n)
    del wbuf
    if bad:
        raise ValueError("witness does not satisfy R1CS (non-zero remainder)")
    log_n = n.bit_length() - 1
    g = fr_to_mont(FR_GENERATOR)
    for x in (a, b, c):
        ctx.ntt(x, n, x, log_n, True)                       # evals -> coefficients
        ctx.ntt(x, n, x, log_n, False, shift_mont=g)        # coefficients -> coset evals
    ctx.fr_mul(a, b, wk.tmp.ptr, n)
    ctx.fr_sub(wk.tmp.ptr, c, wk.tmp.ptr, n)
    z_inv = fr_inv((pow(FR_GENERATOR, n, FR_MODULUS) - 1) % FR_MODULUS)
    ctx.fr_scale(wk.tmp.ptr, fr_to_mont(z_inv), wk.tmp.ptr, n)
    ctx.ntt(wk.tmp.ptr, n, S + 32 * nv, log_n, True, shift_mont=g)   # h coefficients (h[n-1] = 0)
    na = len(pk.aux_ids)
    if na:
        ctx.fr_gather(S, pk.aux_dev.ptr, na, S + 32 * (nv + n))


def _msm_entries(pk: ProvingKey, wk: _Work, slots: Sequence[int]):
    n, nv, na = pk.n, pk.n_vars, len(pk.aux_ids)
    qa, qb, qh, ql = pk.query_offsets
    offs, sizes, boffs = [], [], []
    for sl in slots:
        base = wk.seg * sl
        offs += [base, base, base + nv, base + nv + n]
        sizes += [nv, nv, n - 1, na]
        boffs += [qa, qb, qh, ql]
    return offs, sizes, boffs


def _assemble(pk: ProvingKey, a_lin, b_lin_g1, h_acc, l_acc, b_lin_g2, r: int, s: int) -> Proof:
    vk = pk.vk
    a = g1_add(vk.alpha_g1, g1_add(a_lin, g1_mul(pk.delta_g1, r)))
    b = G2.g2_add(vk.beta_g2, G2.g2_add(b_lin_g2, G2.g2_mul(vk.delta_g2, s)))
    sa = g1_mul(g1_add(vk.alpha_g1, a_lin), s)
    rb = g1_mul(g1_add(pk.beta_g1, b_lin_g1), r)
    rs_delta = g1_mul(pk.delta_g1, r * s % FR_MODULUS)
    c = g1_add(g1_add(l_acc, h_acc), g1_add(g1_add(sa, rb), rs_delta))
    return Proof(a=_norm(a), b=b, c=_norm(c))


def _norm(p):
    from ..curve import from_affine, normalize
    return from_affine(normalize(p))


def check_witness(pk: ProvingKey, w_words: np.ndarray) -> None:
    """A witness is (n_vars, 4) canonical words; anything else is refused
    before any host or device work reads it."""
    shape = getattr(w_words, "shape", None)
    if shape is None or len(shape) != 2 or shape[1] != 4 or shape[0] != pk.n_vars:
        raise ValueError(f"witness must have shape ({pk.n_vars}, 4), got {shape}")


def g2_part(pk: ProvingKey, w_words: np.ndarray, threads: int = 0):
    """b_query_g2 . w on host threads (the reference's MSM_G2 on CPU)."""
    return G2.g2_msm_words(pk.b_query_g2, w_words, pk.n_vars, threads=threads)


def g2_device_set(pk: ProvingKey):
    """The key's b_query_g2 as a device G2 base set, made on the key's
    context at the first call and kept until pk.free().  Creation ends with
    its stream synchronised, so the set serves MSMs on any context of the
    same device."""
    with _G2_SET_LOCK:
        if pk.b_query_g2_dev is None:
            pk.b_query_g2_dev = pk.ctx.msm_g2_bases_create(pk.b_query_g2.words)
        return pk.b_query_g2_dev


_G2_SET_LOCK = threading.Lock()


def device_batch(pk: ProvingKey, witnesses: Sequence[np.ndarray], work: _Work = None,
                 g2: bool = False) -> List[tuple]:
    """Device part of several proofs of one key: per proof the H pipeline,
    then one fixed-base batch for all their G1 MSMs.  Runs on `work`'s
    context (default: the key's own workspace on the key's context), so
    batches of different keys can run concurrently from different host
    threads on different contexts.  Returns per proof (a_lin, b_lin_g1,
    h_acc, l_acc); with `g2` set, per proof ((a_lin, b_lin_g1, h_acc,
    l_acc), b_lin_g2), the G2 MSM run on the device over the witness
    segment the G1 MSMs read."""
    k = len(witnesses)
    if k == 0:
        return []
    wk = work if work is not None else _work(pk, k)
    if not wk.fits(pk, k, wk.ctx):
        raise ValueError("workspace too small for this batch")
    for i, w in enumerate(witnesses):
        _device_part(pk, wk, i, w)
    offs, sizes, boffs = _msm_entries(pk, wk, range(k))
    res = wk.ctx.msm_g1_fixed_batch(pk.fixed, wk.S.ptr, offs, sizes, boffs)
    out = [tuple(from_jacobian_mont(*res[4 * i + j]) for j in range(4)) for i in range(k)]
    if g2:
        gset = g2_device_set(pk)
        for i in range(k):
            words = wk.ctx.msm_g2(gset, wk.S.ptr + 32 * wk.seg * i, pk.n_vars)
            out[i] = (out[i], G2.point_from_words(words))
    return out


def assemble(pk: ProvingKey, g1_parts: tuple, b_lin_g2, r: int, s: int) -> Proof:
    a_lin, b_lin_g1, h_acc, l_acc = g1_parts
    return _assemble(pk, a_lin, b_lin_g1, h_acc, l_acc, b_lin_g2, r, s)


def prove_many(pk: ProvingKey, witnesses: Sequence[np.ndarray], rs: Sequence[Tuple[int, int]],
               g2_threads: int = 0, g2_device: bool = False) -> List[Proof]:
    """Proofs for several witnesses of one key; their G1 MSMs share one
    fixed-base batch launch and their G2 MSMs run on host threads while the
    device works, or on the device after the G1 batch (g2_device)."""
    k = len(witnesses)
    if k == 0:
        return []
    for w in witnesses:
        check_witness(pk, w)
    if g2_device:
        parts = device_batch(pk, witnesses, g2=True)
        return [assemble(pk, g1, g2, *rs[i]) for i, (g1, g2) in enumerate(parts)]
    g2_res: List = [None] * k

    def g2_job():
        for i, w in enumerate(witnesses):
            g2_res[i] = g2_part(pk, w, g2_threads)

    t = threading.Thread(target=g2_job)
    t.start()
    try:
        parts = device_batch(pk, witnesses)
    finally:
        t.join()
    return [assemble(pk, parts[i], g2_res[i], *rs[i]) for i in range(k)]


def prove(pk: ProvingKey, w_words: np.ndarray, r: int, s: int, g2_threads: int = 0,
          g2_device: bool = False) -> Proof:
    return prove_many(pk, [w_words], [(r, s)], g2_threads=g2_threads, g2_device=g2_device)[0]
