"""BN254 G1 group law (oracle; test infrastructure).

The reference delegates to py_ecc.optimized_bn128 (crypto/ecc/bn254.py:11-28:
g1_add / g1_neg / g1_mul / g1_eq by normalisation).  py_ecc is not
importable here, so the group law is restated from its definition on
y^2 = x^3 + 3 over Fq in Jacobian coordinates (x = X/Z^2, y = Y/Z^3,
identity Z = 0); equality and output compare normalised affine points, as
g1_eq does.  Affine points are (x, y) tuples of canonical ints, the identity
is None.

Pinned by: the generator (1, 2) is on the curve and r * G = identity;
g1_add agrees with the textbook affine chord-and-tangent law on random
points (including P + P and P + (-P)); scalar multiplication is
distributive ([a]G + [b]G = [a+b]G); the 16 SRS points of the reference's
PLONK key are on the curve and its 8 commitments equal the naive MSM.
"""
from __future__ import annotations

from .field import Q_MOD, R_MOD

Q = Q_MOD
B = 3
G1_GENERATOR = (1, 2)


def is_on_curve(p) -> bool:
    if p is None:
        return True
    x, y = p
    return (y * y - x * x * x - B) % Q == 0


# ---- affine (textbook) law, used to pin the Jacobian one
def aff_add(p, q):
    if p is None:
        return q
    if q is None:
        return p
    x1, y1 = p
    x2, y2 = q
    if x1 == x2:
        if (y1 + y2) % Q == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, Q) % Q
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, Q) % Q
    x3 = (lam * lam - x1 - x2) % Q
    y3 = (lam * (x1 - x3) - y1) % Q
    return (x3, y3)


def aff_neg(p):
    if p is None:
        return None
    return (p[0], (-p[1]) % Q)


# ---- Jacobian
INF = (1, 1, 0)


def to_jac(p):
    if p is None:
        return INF
    return (p[0], p[1], 1)


def normalize(P):
    X, Y, Z = P
    if Z % Q == 0:
        return None
    zi = pow(Z, -1, Q)
    zi2 = zi * zi % Q
    return (X * zi2 % Q, Y * zi2 * zi % Q)


def jac_double(P):
    X, Y, Z = P
    if Z == 0 or Y == 0:
        return INF
    A = X * X % Q
    Bb = Y * Y % Q
    C = Bb * Bb % Q
    D = 2 * ((X + Bb) * (X + Bb) - A - C) % Q
    E = 3 * A % Q
    F = E * E % Q
    X3 = (F - 2 * D) % Q
    Y3 = (E * (D - X3) - 8 * C) % Q
    Z3 = 2 * Y * Z % Q
    return (X3, Y3, Z3)


def jac_add(P, Qp):
    X1, Y1, Z1 = P
    X2, Y2, Z2 = Qp
    if Z1 == 0:
        return Qp
    if Z2 == 0:
        return P
    Z1Z1 = Z1 * Z1 % Q
    Z2Z2 = Z2 * Z2 % Q
    U1 = X1 * Z2Z2 % Q
    U2 = X2 * Z1Z1 % Q
    S1 = Y1 * Z2 * Z2Z2 % Q
    S2 = Y2 * Z1 * Z1Z1 % Q
    if U1 == U2:
        if S1 == S2:
            return jac_double(P)
        return INF
    H = (U2 - U1) % Q
    I = 4 * H * H % Q
    J = H * I % Q
    rr = 2 * (S2 - S1) % Q
    V = U1 * I % Q
    X3 = (rr * rr - J - 2 * V) % Q
    Y3 = (rr * (V - X3) - 2 * S1 * J) % Q
    Z3 = ((Z1 + Z2) * (Z1 + Z2) - Z1Z1 - Z2Z2) * H % Q
    return (X3, Y3, Z3)


def jac_neg(P):
    return (P[0], (-P[1]) % Q, P[2])


def jac_mul(P, k: int):
    """[k]P by double-and-add (py_ecc.multiply semantics: k is not reduced,
    but the group has order r so the result equals [k mod r]P)."""
    k = int(k)
    if k < 0:
        return jac_mul(jac_neg(P), -k)
    acc = INF
    for bit in bin(k)[2:] if k else "":
        acc = jac_double(acc)
        if bit == "1":
            acc = jac_add(acc, P)
    return acc


def g1_mul_affine(p, k: int):
    return normalize(jac_mul(to_jac(p), k))


def g1_add_affine(p, q):
    return normalize(jac_add(to_jac(p), to_jac(q)))
