"""CPU MSM oracle (test infrastructure).

  default_window_bits   runtime/kernels/cpu/kernels.py:557-560 and
                        runtime/metal/msm.py:72,191: c = max(4, min(16, bitlen(n) - 2))
  msm_naive_g1          crypto/msm/cpu_ref.py:8-12
  msm_pippenger         crypto/msm/pippenger.py:10-64 (unsigned c-bit windows
                        over bitlen(r) = 254 bits; dense running sum for
                        2^c <= 1024 buckets, sparse gap form above)
  encode_signed_digits  runtime/metal/msm.py:145-177
  csr_buckets           runtime/metal/ecc_source.py:437-496 (histogram,
                        exclusive scan, scatter of sign<<31 | index) with a
                        stable order instead of the atomic one
  msm_signed_buckets    runtime/metal/ecc_source.py:499-566 and
                        runtime/metal/msm.py:322-332 (bucket sums, running-sum
                        window reduction, host Horner combination)
  msm_closed_form       [sum_i s_i k_i mod r] G for bases B_i = [k_i] G

All points here are affine tuples (or None for the identity); the internal
accumulation is Jacobian (oracle/curve.py).

Pinned by: naive == Pippenger == signed-bucket == closed form on every
size n in {1, 2, 3, 8, 31, 32, 63, 64, 127} and window 7/8
(tests/correctness/cpu/test_msm_pippenger.py:11-27 strategy) plus larger
random cases; the 8 commitments of the reference's PLONK key
(tests/golden/plonk_pk_n16.json); signed-digit invariants
sum_k d_k 2^(ck) = s, |d_k| <= 2^(c-1), zero final carry; CSR against a
host stable sort.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

from . import curve as C
from .field import R_MOD


def default_window_bits(n: int) -> int:
    return max(4, min(16, int(n).bit_length() - 2))


def msm_naive_g1(points: Sequence, scalars: Sequence[int]):
    acc = C.INF
    for p, s in zip(points, scalars):
        acc = C.jac_add(acc, C.jac_mul(C.to_jac(p), int(s)))
    return C.normalize(acc)


def msm_pippenger(points: Sequence, scalars: Sequence[int], window_bits: int = 16):
    if len(points) != len(scalars):
        raise ValueError("points and scalars length mismatch")
    if len(points) == 0:
        return None
    w = int(window_bits)
    if w <= 0 or w > 20:
        raise ValueError("invalid window_bits")
    max_bits = R_MOD.bit_length()
    nb = 1 << w
    pts = [C.to_jac(p) for p in points]
    ss = [int(s) % R_MOD for s in scalars]
    acc = C.INF
    for k in reversed(range(0, max_bits, w)):
        window_sum = C.INF
        if nb <= 1024:
            buckets = [C.INF] * nb
            for p, s in zip(pts, ss):
                d = (s >> k) & (nb - 1)
                if d:
                    buckets[d] = C.jac_add(buckets[d], p)
            running = C.INF
            for d in range(nb - 1, 0, -1):
                running = C.jac_add(running, buckets[d])
                window_sum = C.jac_add(window_sum, running)
        else:
            b2 = {}
            for p, s in zip(pts, ss):
                d = (s >> k) & (nb - 1)
                if d:
                    b2[d] = C.jac_add(b2[d], p) if d in b2 else p
            running = C.INF
            prev = nb - 1
            for d in sorted(b2, reverse=True):
                gap = prev - d - 1
                if gap > 0 and running[2] != 0:
                    window_sum = C.jac_add(window_sum, C.jac_mul(running, gap))
                running = C.jac_add(running, b2[d])
                window_sum = C.jac_add(window_sum, running)
                prev = d
            tail = prev - 1
            if tail > 0 and running[2] != 0:
                window_sum = C.jac_add(window_sum, C.jac_mul(running, tail))
        acc = C.jac_mul(acc, 1 << w)
        acc = C.jac_add(acc, window_sum)
    return C.normalize(acc)


def window_count(c: int) -> int:
    """runtime/metal/msm.py:194: ceil((254 + 1) / c)."""
    return (R_MOD.bit_length() + 1 + c - 1) // c


def encode_signed_digits(scalars: Sequence[int], c: int, W: int) -> List[List[int]]:
    """Per scalar, W signed digits d_k in [-2^(c-1), 2^(c-1) - 1] (value
    before sign fix in [0, 2^c]; >= 2^(c-1) becomes v - 2^c with carry)."""
    half = 1 << (c - 1)
    mask = (1 << c) - 1
    out = []
    for s in scalars:
        t = int(s)
        carry = 0
        row = []
        for _ in range(W):
            v = (t & mask) + carry
            t >>= c
            if v >= half:
                row.append(v - (1 << c))
                carry = 1
            else:
                row.append(v)
                carry = 0
        out.append(row)
    return out


def csr_buckets(digits: List[List[int]], W: int, c: int) -> Tuple[List[int], List[int]]:
    """row_ptr over W * (2^(c-1) + 1) buckets (window-major) and the entries
    sign<<31 | index, stable by index within a bucket."""
    B = (1 << (c - 1)) + 1
    counts = [0] * (W * B)
    for row in digits:
        for w, d in enumerate(row):
            if d:
                counts[w * B + abs(d)] += 1
    row_ptr = [0] * (W * B + 1)
    for i in range(W * B):
        row_ptr[i + 1] = row_ptr[i] + counts[i]
    cur = list(row_ptr[:-1])
    entries = [0] * row_ptr[-1]
    for w in range(W):
        for i, row in enumerate(digits):
            d = row[w]
            if d:
                key = w * B + abs(d)
                entries[cur[key]] = i | ((1 << 31) if d < 0 else 0)
                cur[key] += 1
    return row_ptr, entries


def msm_signed_buckets(points: Sequence, scalars: Sequence[int], c: int):
    n = len(points)
    if n == 0:
        return None
    W = window_count(c)
    B = (1 << (c - 1)) + 1
    digits = encode_signed_digits([int(s) % R_MOD for s in scalars], c, W)
    row_ptr, entries = csr_buckets(digits, W, c)
    pts = [C.to_jac(p) for p in points]
    window_sums = []
    for w in range(W):
        buckets = [C.INF] * B
        for b in range(1, B):
            key = w * B + b
            acc = C.INF
            for e in entries[row_ptr[key]:row_ptr[key + 1]]:
                p = pts[e & 0x7FFFFFFF]
                acc = C.jac_add(acc, C.jac_neg(p) if e >> 31 else p)
            buckets[b] = acc
        running = C.INF
        ws = C.INF
        for b in range(B - 1, 0, -1):
            running = C.jac_add(running, buckets[b])
            ws = C.jac_add(ws, running)
        window_sums.append(ws)
    acc = C.INF
    for w in reversed(range(W)):
        if w != W - 1:
            acc = C.jac_mul(acc, 1 << c)
        acc = C.jac_add(acc, window_sums[w])
    return C.normalize(acc)


def msm_closed_form(ks: Sequence[int], ss: Sequence[int]):
    t = 0
    for k, s in zip(ks, ss):
        t += int(k) * int(s)
    return C.g1_mul_affine(C.G1_GENERATOR, t % R_MOD)
