"""Host-side BN254 G1 helpers for the operator layer.

Host G1 points use the reference's py_ecc.optimized_bn128 convention
(crypto/ecc/bn254.py:7-54): homogeneous projective (X, Y, Z) tuples with
x = X/Z, y = Y/Z; the identity is (1, 1, 0).  Here the coordinates are plain
ints mod q.  ``g1_eq`` compares normalised affine points
(crypto/ecc/bn254.py:27-28).  Device results arrive as Jacobian Montgomery
(X, Y, Z) (runtime/metal/msm.py:36-63) and are normalised to (x, y, 1).
"""
from __future__ import annotations

from typing import Optional, Tuple

from .field import FQ_MODULUS, fq_from_mont, fq_to_mont

Q = FQ_MODULUS
G1Point = Tuple[int, int, int]
G1_GENERATOR: G1Point = (1, 2, 1)
G1_ZERO: G1Point = (1, 1, 0)


def _coord(v) -> int:
    # accept plain ints or objects exposing an ``n`` attribute (py_ecc FQ)
    return int(v.n) if hasattr(v, "n") else int(v)


def normalize(p) -> Optional[Tuple[int, int]]:
    """Affine (x, y) of a homogeneous point, None for the identity."""
    X, Y, Z = (_coord(c) % Q for c in p)
    if Z == 0:
        return None
    zi = pow(Z, -1, Q)
    return (X * zi % Q, Y * zi % Q)


def g1_eq(p, q) -> bool:
    return normalize(p) == normalize(q)


def is_on_curve(p) -> bool:
    a = normalize(p)
    if a is None:
        return True
    x, y = a
    return (y * y - x * x * x - 3) % Q == 0


def from_affine(a: Optional[Tuple[int, int]]) -> G1Point:
    return G1_ZERO if a is None else (a[0] % Q, a[1] % Q, 1)


def from_jacobian_mont(X: int, Y: int, Z: int) -> G1Point:
    """Device result (Jacobian, Montgomery) -> normalised host point."""
    if Z == 0:
        return G1_ZERO
    x, y, z = fq_from_mont(X), fq_from_mont(Y), fq_from_mont(Z)
    zi = pow(z, -1, Q)
    zi2 = zi * zi % Q
    return (x * zi2 % Q, y * zi2 * zi % Q, 1)


def affine_mont_bytes(points) -> bytes:
    """Pack host points as 64-byte affine Montgomery (x, y); the identity is
    packed as x = y = 0 (replaces runtime/metal/msm.py:13-34)."""
    out = bytearray()
    for p in points:
        a = normalize(p)
        if a is None:
            out += bytes(64)
        else:
            out += fq_to_mont(a[0]).to_bytes(32, "little")
            out += fq_to_mont(a[1]).to_bytes(32, "little")
    return bytes(out)


# ---- group law on homogeneous coordinates (add-1998-cmo-2 / dbl-2007-bl
# shapes for a = 0), used for the few host-side point operations of the
# operator layer (combining partial results, scalar multiples of single points)
def g1_double(p: G1Point) -> G1Point:
    X, Y, Z = (c % Q for c in p)
    if Z == 0 or Y == 0:
        return G1_ZERO
    W = 3 * X * X % Q
    S = Y * Z % Q
    B = X * Y * S % Q
    H = (W * W - 8 * B) % Q
    S2 = S * S % Q
    X3 = 2 * H * S % Q
    Y3 = (W * (4 * B - H) - 8 * Y * Y * S2) % Q
    Z3 = 8 * S * S2 % Q
    return (X3, Y3, Z3)


def g1_add(p: G1Point, q: G1Point) -> G1Point:
    X1, Y1, Z1 = (c % Q for c in p)
    X2, Y2, Z2 = (c % Q for c in q)
    if Z1 == 0:
        return (X2, Y2, Z2)
    if Z2 == 0:
        return (X1, Y1, Z1)
    U1 = Y2 * Z1 % Q
    U2 = Y1 * Z2 % Q
    V1 = X2 * Z1 % Q
    V2 = X1 * Z2 % Q
    if V1 == V2:
        if U1 == U2:
            return g1_double((X1, Y1, Z1))
        return G1_ZERO
    U = (U1 - U2) % Q
    V = (V1 - V2) % Q
    V_sq = V * V % Q
    V_sq_V2 = V_sq * V2 % Q
    V_cu = V * V_sq % Q
    W = Z1 * Z2 % Q
    A = (U * U * W - V_cu - 2 * V_sq_V2) % Q
    return (V * A % Q, (U * (V_sq_V2 - A) - V_cu * U2) % Q, V_cu * W % Q)


def g1_neg(p: G1Point) -> G1Point:
    X, Y, Z = p
    return (X % Q, (-Y) % Q, Z % Q)


def g1_mul(p: G1Point, k: int) -> G1Point:
    k = int(k)
    if k < 0:
        return g1_mul(g1_neg(p), -k)
    acc = G1_ZERO
    for bit in bin(k)[2:] if k else "":
        acc = g1_double(acc)
        if bit == "1":
            acc = g1_add(acc, p)
    return acc


def g1_mul_native(p: G1Point, k: int) -> G1Point:
    """[k] P through the library's host multiplier (one call, no Python
    double-and-add); same result as g1_mul."""
    import numpy as np
    from ._native import load_library
    from .field import FR_MODULUS, ints_to_words, words_to_ints
    a = normalize(p)
    src = np.zeros(8, dtype="<u8")
    if a is not None:
        src[:] = ints_to_words([a[0], a[1]]).reshape(-1)
    kw = ints_to_words([int(k) % FR_MODULUS]).reshape(-1).astype("<u8")
    out = np.zeros(8, dtype="<u8")
    rc = load_library().zk_g1_mul_host(src.ctypes.data, kw.ctypes.data, out.ctypes.data)
    if rc:
        raise ValueError("zk_g1_mul_host failed")
    x, y = words_to_ints(out.reshape(2, 4))
    return from_affine(None if x == 0 and y == 0 else (x, y))


def fixed_base_table(p: G1Point = G1_GENERATOR) -> bytes:
    """T[j][d] = [d * 2^(8j)] p for j < 32, d < 256, packed as 64-byte affine
    Montgomery points (the table zk_g1_fixed_base expects)."""
    rows = []
    base = p
    for _ in range(32):
        row = [G1_ZERO]
        acc = G1_ZERO
        for _ in range(255):
            acc = g1_add(acc, base)
            row.append(acc)
        rows.extend(row)
        base = g1_add(acc, base)  # 256 * base
    return affine_mont_bytes(rows)
