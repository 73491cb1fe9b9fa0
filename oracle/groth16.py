"""Groth16 oracle (test infrastructure), restating the reference:

  setup   protocols/groth16/setup.py:25-121: per variable, interpolate the
          A/B/C column over the domain with an inverse NTT and evaluate at
          tau (a_tau, b_tau, c_tau); a_query = [a_tau]G1, b_query = [b_tau]
          G2 / G1, ic / l_query = [(beta a + alpha b + c)/gamma or delta]G1,
          h_query[k] = [tau^k t(tau)/delta]G1 for k < n - 1.
  prove   protocols/groth16/prove.py:30-128 with the QAP of qap.py:53-101:
          iNTT of A w, B w, C w; H = (A B - C) / (X^n - 1) via NTT product
          (crypto/poly/ntt.py:149-162) and the long division of
          crypto/poly/fast.py:8-37; MSMs by Pippenger; C = l + h + s(alpha +
          a_lin) + r(beta1 + b_lin_g1) + r s delta1.
  check_in_exponent: the verification equation e(A,B) = e(alpha,beta)
          e(vk_x,gamma) e(C,delta) in the exponent, from the trapdoor and the
          witness, plus A = [a]G1, B = [b]G2, C = [c]G1 — a pairing-free
          check of a proof whose discrete logs the test knows.

Pinned by: check_in_exponent on its own proofs (an independent derivation
of the dlogs), H (x) Z_H == A B - C identically, R1CS satisfaction; small
circuits only.
"""
from __future__ import annotations

from typing import Dict, List, Sequence

from . import curve as C1
from . import curve2 as C2
from . import msm as M
from . import ntt as N
from .field import R_MOD


def _col(rows: List[Dict[int, int]], var: int) -> List[int]:
    return [int(r.get(var, 0)) % R_MOD for r in rows]


def poly_mul(a: Sequence[int], b: Sequence[int]) -> List[int]:
    """crypto/poly/ntt.py:149-162"""
    if not a or not b:
        return []
    need = len(a) + len(b) - 1
    n = 1 << (need - 1).bit_length()
    w = N.omega_for_size(n)
    ae = N.evals_from_coeffs_on_roots(a, n, w)
    be = N.evals_from_coeffs_on_roots(b, n, w)
    out = N.coeffs_from_evals_on_roots([x * y % R_MOD for x, y in zip(ae, be)], w)[:need]
    while out and out[-1] == 0:
        out.pop()
    return out


def div_xn_minus_1(num: Sequence[int], n: int):
    """crypto/poly/fast.py:8-37"""
    a = [int(c) % R_MOD for c in num]
    while a and a[-1] == 0:
        a.pop()
    if not a:
        return [], []
    if len(a) <= n:
        return [], a
    m = len(a) - 1
    qd = m - n
    q = [0] * (qd + 1)
    for k in range(m, n - 1, -1):
        qk = q[k] if k <= qd else 0
        q[k - n] = (a[k] + qk) % R_MOD
    r = [0] * n
    for i in range(n):
        qi = q[i] if i <= qd else 0
        r[i] = ((a[i] if i < len(a) else 0) + qi) % R_MOD
    while q and q[-1] == 0:
        q.pop()
    while r and r[-1] == 0:
        r.pop()
    return q, r


def setup(A, B, C, n: int, n_vars: int, public_ids, aux_ids, trapdoor):
    tau, alpha, beta, gamma, delta = trapdoor
    w = N.omega_for_size(n)
    inv_g, inv_d = pow(gamma, -1, R_MOD), pow(delta, -1, R_MOD)
    G1, G2 = C1.G1_GENERATOR, C2.G2_GENERATOR
    a_tau, b_tau, k_tau = [], [], []
    for var in range(n_vars):
        ac = N.coeffs_from_evals_on_roots(_col(A, var), w)
        bc = N.coeffs_from_evals_on_roots(_col(B, var), w)
        cc = N.coeffs_from_evals_on_roots(_col(C, var), w)
        at, bt, ctau = N.poly_eval(ac, tau), N.poly_eval(bc, tau), N.poly_eval(cc, tau)
        a_tau.append(at)
        b_tau.append(bt)
        k_tau.append((beta * at + alpha * bt + ctau) % R_MOD)
    t_tau = (pow(tau, n, R_MOD) - 1) % R_MOD
    pk = {
        "a_query": [C1.g1_mul_affine(G1, x) for x in a_tau],
        "b_query_g1": [C1.g1_mul_affine(G1, x) for x in b_tau],
        "b_query_g2": [C2.mul(G2, x) for x in b_tau],
        "h_query": [C1.g1_mul_affine(G1, pow(tau, k, R_MOD) * t_tau * inv_d % R_MOD) for k in range(n - 1)],
        "l_query": [C1.g1_mul_affine(G1, k_tau[i] * inv_d % R_MOD) for i in aux_ids],
        "ic": [C1.g1_mul_affine(G1, k_tau[i] * inv_g % R_MOD) for i in public_ids],
        "alpha_g1": C1.g1_mul_affine(G1, alpha), "beta_g1": C1.g1_mul_affine(G1, beta),
        "beta_g2": C2.mul(G2, beta), "gamma_g2": C2.mul(G2, gamma), "delta_g1": C1.g1_mul_affine(G1, delta),
        "delta_g2": C2.mul(G2, delta), "aux_ids": list(aux_ids), "n": n,
        "trapdoor": trapdoor, "a_tau": a_tau, "b_tau": b_tau, "k_tau": k_tau, "public_ids": list(public_ids),
    }
    return pk


def _row(row, w):
    return sum(int(c) * w[v] for v, c in row.items()) % R_MOD


def compute_h(A, B, C, w, n):
    """qap.py:53-101 (raises ValueError on a non-zero remainder)."""
    om = N.omega_for_size(n)
    a = N.coeffs_from_evals_on_roots([_row(r, w) for r in A], om)
    b = N.coeffs_from_evals_on_roots([_row(r, w) for r in B], om)
    c = N.coeffs_from_evals_on_roots([_row(r, w) for r in C], om)
    p = poly_mul(a, b)
    p = [((p[i] if i < len(p) else 0) - (c[i] if i < len(c) else 0)) % R_MOD for i in range(max(len(p), len(c)))]
    q, rem = div_xn_minus_1(p, n)
    if rem:
        raise ValueError("witness does not satisfy R1CS (non-zero remainder)")
    h = list(q) + [0] * (n - 1 - len(q))
    return h[: n - 1]


def prove(pk, A, B, C, w: Sequence[int], r: int, s: int):
    n = pk["n"]
    a_lin = M.msm_pippenger(pk["a_query"], w, 4)
    b_lin_g1 = M.msm_pippenger(pk["b_query_g1"], w, 4)
    b_lin_g2 = C2.msm_naive(pk["b_query_g2"], w)
    h = compute_h(A, B, C, w, n)
    h_acc = M.msm_pippenger(pk["h_query"], h, 4) if h else None
    l_acc = M.msm_pippenger(pk["l_query"], [w[i] for i in pk["aux_ids"]], 4) if pk["aux_ids"] else None
    add, mul = C1.aff_add, C1.g1_mul_affine
    a = add(pk["alpha_g1"], add(a_lin, mul(pk["delta_g1"], r)))
    b = C2.add(pk["beta_g2"], C2.add(b_lin_g2, C2.mul(pk["delta_g2"], s)))
    sa = mul(add(pk["alpha_g1"], a_lin), s)
    rb = mul(add(pk["beta_g1"], b_lin_g1), r)
    rsd = mul(pk["delta_g1"], r * s % R_MOD)
    c = add(add(l_acc, h_acc), add(add(sa, rb), rsd))
    return {"a": a, "b": b, "c": c, "h": h}


def exponent_key(A, B, C, n: int, n_vars: int, public_ids, trapdoor):
    """The part of `setup`'s key that check_in_exponent reads (the QAP
    columns at tau), without the group elements: u_i(tau) = sum_j A[j][i]
    L_j(tau) with the Lagrange basis L_j(tau) = w^j (tau^n - 1) / (n (tau -
    w^j)) of the domain, since u_i interpolates column i on the domain.
    Pinned against `setup` (iNTT then Horner) on small circuits."""
    tau, alpha, beta, gamma, delta = trapdoor
    w = N.omega_for_size(n)
    zt = (pow(tau, n, R_MOD) - 1) % R_MOD
    n_inv = pow(n, -1, R_MOD)
    lag, x = [], 1
    for _ in range(n):
        lag.append(x * zt % R_MOD * n_inv % R_MOD * pow((tau - x) % R_MOD, -1, R_MOD) % R_MOD)
        x = x * w % R_MOD
    cols = []
    for rows in (A, B, C):
        acc = [0] * n_vars
        for j, row in enumerate(rows):
            for var, c in row.items():
                acc[var] = (acc[var] + int(c) * lag[j]) % R_MOD
        cols.append(acc)
    a_tau, b_tau, c_tau = cols
    k_tau = [(beta * a + alpha * b + c) % R_MOD for a, b, c in zip(a_tau, b_tau, c_tau)]
    return {"trapdoor": trapdoor, "a_tau": a_tau, "b_tau": b_tau, "k_tau": k_tau, "public_ids": list(public_ids),
            "n": n}


def check_in_exponent(pk, w: Sequence[int], r: int, s: int, proof, n: int) -> bool:
    tau, alpha, beta, gamma, delta = pk["trapdoor"]
    a_tau, b_tau, k_tau = pk["a_tau"], pk["b_tau"], pk["k_tau"]
    a = (alpha + sum(x * y for x, y in zip(w, a_tau)) + r * delta) % R_MOD
    b = (beta + sum(x * y for x, y in zip(w, b_tau)) + s * delta) % R_MOD
    x_pub = sum(w[i] * k_tau[i] for i in pk["public_ids"]) % R_MOD          # gamma * dlog(vk_x)
    # c from the verification equation: a b = alpha beta + x_pub + c delta
    c = (a * b - alpha * beta - x_pub) * pow(delta, -1, R_MOD) % R_MOD
    G1, G2 = C1.G1_GENERATOR, C2.G2_GENERATOR
    return (C1.g1_mul_affine(G1, a) == proof["a"] and C2.mul(G2, b) == proof["b"]
            and C1.g1_mul_affine(G1, c) == proof["c"])
