"""PLONK quotient oracle (test infrastructure).

  batch_inv        crypto/field/batch.py:8-38  (0 -> 0)
  quotient_evals   runtime/kernels/cpu/kernels.py:607-694 (PLONK_T_QUOTIENT_EVALS)
  quotient         protocols/plonk/prove.py:300-424 (_compute_quotient: 15 coset
                   NTTs at m = 4n, the fused numerator, BATCH_INV, point-wise
                   product, coset iNTT, degree check, split into t1, t2, t3)

The shift is an argument (the reference draws it with `secrets`,
prove.py:323-327, rejecting shift^n in {1, u, u^2, u^3}, u = omega_m^n);
L1's coefficients come from the inverse NTT of [1, 0, ..., 0] as in
prove.py:329-330.

Pinned by: batch_inv against pow(x, -1) with zeros; the quotient pipeline
against an independent evaluation — Horner evaluation of every input
polynomial at each coset point, the numerator formula, division by
x^n - 1 and the O(m^2) inverse DFT on the coset; and a multiplication-gate
instance (qm = 1, qo = -1, identity permutation, z = 1) whose quotient t
must satisfy t(X) Z_H(X) = a(X) b(X) - c(X), with a wrong witness raising
"quotient degree too large".
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

from .field import R_MOD, fr_inv
from .ntt import (coeffs_from_evals_on_coset, coeffs_from_evals_on_roots, evals_from_coeffs_on_coset,
                  omega_for_size)

Q_INPUTS = ("a", "b", "c", "z", "pi", "ql", "qr", "qm", "qo", "qc", "s1", "s2", "s3", "zshift", "l1")


def batch_inv(xs: Sequence[int]) -> List[int]:
    """Montgomery's simultaneous inversion, zeros map to zero (batch.py:8-38)."""
    n = len(xs)
    if n == 0:
        return []
    prefix = [1] * (n + 1)
    acc = 1
    for i, x in enumerate(xs):
        xi = int(x) % R_MOD
        if xi:
            acc = acc * xi % R_MOD
        prefix[i + 1] = acc
    out = [0] * n
    suffix = fr_inv(acc)
    for i in range(n - 1, -1, -1):
        xi = int(xs[i]) % R_MOD
        if xi:
            out[i] = prefix[i] * suffix % R_MOD
            suffix = suffix * xi % R_MOD
    return out


def quotient_evals(ext: Dict[str, Sequence[int]], *, n: int, m: int, omega_m: int, shift: int, shift_n: int,
                   alpha: int, beta: int, gamma: int, k1: int, k2: int) -> Tuple[List[int], List[int]]:
    """Numerator and Z_H at x_i = shift omega_m^i (kernels.py:607-694)."""
    if not all(len(ext[k]) == m for k in Q_INPUTS):
        raise ValueError("eval length mismatch")
    alpha, beta, gamma, k1, k2 = (int(v) % R_MOD for v in (alpha, beta, gamma, k1, k2))
    alpha2 = alpha * alpha % R_MOD
    u = pow(int(omega_m) % R_MOD, n, R_MOD)
    x = int(shift) % R_MOD
    u_pow = 1
    num, zh = [], []
    for i in range(m):
        a, b, c, z = (int(ext[k][i]) % R_MOD for k in ("a", "b", "c", "z"))
        gate = (int(ext["ql"][i]) * a + int(ext["qr"][i]) * b + int(ext["qm"][i]) * a % R_MOD * b
                + int(ext["qo"][i]) * c + int(ext["qc"][i]) + int(ext["pi"][i])) % R_MOD
        left = ((a + beta * x + gamma) * (b + beta * k1 % R_MOD * x + gamma) % R_MOD
                * (c + beta * k2 % R_MOD * x + gamma) % R_MOD * z) % R_MOD
        right = ((a + beta * int(ext["s1"][i]) + gamma) * (b + beta * int(ext["s2"][i]) + gamma) % R_MOD
                 * (c + beta * int(ext["s3"][i]) + gamma) % R_MOD * int(ext["zshift"][i])) % R_MOD
        boundary = (z - 1) % R_MOD * (int(ext["l1"][i]) % R_MOD) % R_MOD
        num.append((gate + alpha * (left - right) + alpha2 * boundary) % R_MOD)
        zh.append((int(shift_n) * u_pow - 1) % R_MOD)
        u_pow = u_pow * u % R_MOD
        x = x * int(omega_m) % R_MOD
    return num, zh


def extended_inputs(coeffs: Dict[str, Sequence[int]], *, n: int, omega: int, shift: int) -> Dict[str, List[int]]:
    """The 15 coset evaluation vectors of prove.py:366-381 (l1 from the iNTT
    of [1, 0, ...], z(omega X) as the coset at shift * omega)."""
    m = 4 * n
    omega_m = omega_for_size(m)
    l1 = coeffs_from_evals_on_roots([1] + [0] * (n - 1), omega)
    src = dict(coeffs)
    src["l1"] = l1
    ext = {}
    for k in Q_INPUTS:
        if k == "zshift":
            ext[k] = evals_from_coeffs_on_coset(coeffs["z"], n=m, omega=omega_m, shift=shift * omega % R_MOD)
        else:
            ext[k] = evals_from_coeffs_on_coset(src[k], n=m, omega=omega_m, shift=shift)
    return ext


def quotient(coeffs: Dict[str, Sequence[int]], *, n: int, omega: int, k1: int, k2: int, alpha: int, beta: int,
             gamma: int, shift: int) -> Tuple[List[int], List[int], List[int]]:
    """t1, t2, t3 (each n coefficients) of prove.py:300-424 for a given shift;
    raises ValueError("quotient degree too large") as the reference does."""
    m = 4 * n
    omega_m = omega_for_size(m)
    ext = extended_inputs(coeffs, n=n, omega=omega, shift=shift)
    num, zh = quotient_evals(ext, n=n, m=m, omega_m=omega_m, shift=shift, shift_n=pow(shift, n, R_MOD),
                             alpha=alpha, beta=beta, gamma=gamma, k1=k1, k2=k2)
    inv = batch_inv(zh)
    t_ext = [a * b % R_MOD for a, b in zip(num, inv)]
    t_full = coeffs_from_evals_on_coset(t_ext, omega=omega_m, shift=shift)
    for v in t_full[3 * n:]:
        if v % R_MOD != 0:
            raise ValueError("quotient degree too large")
    t = list(t_full[:3 * n]) + [0] * max(0, 3 * n - len(t_full))
    return t[0:n], t[n:2 * n], t[2 * n:3 * n]


def quotient_full(coeffs: Dict[str, Sequence[int]], *, n: int, omega: int, k1: int, k2: int, alpha: int,
                  beta: int, gamma: int, shift: int) -> List[int]:
    """All m coefficients of t before the degree check (for inputs that are
    not a valid PLONK instance)."""
    m = 4 * n
    omega_m = omega_for_size(m)
    ext = extended_inputs(coeffs, n=n, omega=omega, shift=shift)
    num, zh = quotient_evals(ext, n=n, m=m, omega_m=omega_m, shift=shift, shift_n=pow(shift, n, R_MOD),
                             alpha=alpha, beta=beta, gamma=gamma, k1=k1, k2=k2)
    inv = batch_inv(zh)
    return coeffs_from_evals_on_coset([a * b % R_MOD for a, b in zip(num, inv)], omega=omega_m, shift=shift)
