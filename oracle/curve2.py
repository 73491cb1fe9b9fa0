"""BN254 G2 group law (oracle; test infrastructure).

The reference takes G2 from py_ecc.optimized_bn128 (crypto/ecc/bn254.py:31-54).
Restated from the definition: Fq2 = Fq[u]/(u^2 + 1), the twist
E': y^2 = x^3 + 3/(9 + u), affine chord-and-tangent law; points are
((x0, x1), (y0, y1)) or None for the identity.  The generator is py_ecc's
G2 (it equals the reference PLONK key's first G2 SRS point,
tests/golden/plonk_pk_n16.json).

Pinned by: the generator and the key's second G2 SRS point lie on the
twist, r * G2 = identity, distributivity [a]Q + [b]Q = [a+b]Q.
"""
from __future__ import annotations

from .field import Q_MOD, R_MOD

Q = Q_MOD


def f2_add(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def f2_sub(a, b):
    return ((a[0] - b[0]) % Q, (a[1] - b[1]) % Q)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def f2_inv(a):
    t = pow((a[0] * a[0] + a[1] * a[1]) % Q, -1, Q)
    return (a[0] * t % Q, (-a[1]) * t % Q)


def f2_scal(a, k):
    return (a[0] * k % Q, a[1] * k % Q)


B2 = f2_mul((3, 0), f2_inv((9, 1)))

G2_GENERATOR = (
    (10857046999023057135944570762232829481370756359578518086990519993285655852781,
     11559732032986387107991004021392285783925812861821192530917403151452391805634),
    (8495653923123431417604973247489272438418190587263600148770280649306958101930,
     4082367875863433681332203403145435568316851327593401208105741076214120093531),
)


def is_on_twist(p) -> bool:
    if p is None:
        return True
    x, y = p
    return f2_sub(f2_mul(y, y), f2_add(f2_mul(f2_mul(x, x), x), B2)) == (0, 0)


def add(p, q):
    if p is None:
        return q
    if q is None:
        return p
    (x1, y1), (x2, y2) = p, q
    if x1 == x2:
        if f2_add(y1, y2) == (0, 0):
            return None
        lam = f2_mul(f2_scal(f2_mul(x1, x1), 3), f2_inv(f2_scal(y1, 2)))
    else:
        lam = f2_mul(f2_sub(y2, y1), f2_inv(f2_sub(x2, x1)))
    x3 = f2_sub(f2_sub(f2_mul(lam, lam), x1), x2)
    y3 = f2_sub(f2_mul(lam, f2_sub(x1, x3)), y1)
    return (x3, y3)


def neg(p):
    return None if p is None else (p[0], ((-p[1][0]) % Q, (-p[1][1]) % Q))


def mul(p, k: int):
    k = int(k)
    acc = None
    for bit in bin(k)[2:] if k else "":
        acc = add(acc, acc)
        if bit == "1":
            acc = add(acc, p)
    return acc


def msm_naive(points, scalars):
    acc = None
    for p, s in zip(points, scalars):
        acc = add(acc, mul(p, int(s) % R_MOD))
    return acc
