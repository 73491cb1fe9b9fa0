"""CPU NTT oracle (test infrastructure): restates crypto/poly/ntt.py.

  omega_for_size        crypto/poly/ntt.py:13-24
  ntt_inplace           crypto/poly/ntt.py:42-75  (bit reversal + radix-2 DIT)
  intt_inplace          crypto/poly/ntt.py:79-87
  evals_from_coeffs_on_roots   crypto/poly/ntt.py:103-110 (truncate / zero-pad)
  coeffs_from_evals_on_roots   crypto/poly/ntt.py:113-118
  evals_from_coeffs_on_coset   crypto/poly/ntt.py:121-132
  coeffs_from_evals_on_coset   crypto/poly/ntt.py:135-146 (shift 0 rule)
  dft_naive             O(n^2) definition, the closed form used to pin the above

Pinned by: the O(n^2) DFT definition for every n <= 64 and both directions,
the 8 inverse-NTT vectors of the reference's cached PLONK key
(tests/golden/plonk_pk_n16.json), Horner evaluation at shift * omega^i for
the coset forms (tests/correctness/cpu/test_poly_ntt.py:77-86 strategy),
delta -> all-ones and constant -> c*n*delta_0 identities.
"""
from __future__ import annotations

from typing import List, Sequence

from .field import MAX_ORDER_LOG, R_MOD, ROOT_2_28, fr_inv


def omega_for_size(n: int) -> int:
    if n <= 0 or n & (n - 1):
        raise ValueError("n must be a power of two")
    logn = n.bit_length() - 1
    if logn > MAX_ORDER_LOG:
        raise ValueError("domain too large for BN254 root of unity")
    return pow(ROOT_2_28, 1 << (MAX_ORDER_LOG - logn), R_MOD)


def ntt_inplace(a: List[int], omega: int) -> None:
    n = len(a)
    if n == 0:
        return
    if n & (n - 1):
        raise ValueError("length must be power of two")
    j = 0
    for i in range(1, n):
        bit = n >> 1
        while j & bit:
            j ^= bit
            bit >>= 1
        j ^= bit
        if i < j:
            a[i], a[j] = a[j], a[i]
    length = 2
    w0 = int(omega) % R_MOD
    while length <= n:
        wlen = pow(w0, n // length, R_MOD)
        half = length >> 1
        for i in range(0, n, length):
            w = 1
            for k in range(half):
                u = a[i + k] % R_MOD
                v = (a[i + k + half] % R_MOD) * w % R_MOD
                a[i + k] = (u + v) % R_MOD
                a[i + k + half] = (u - v) % R_MOD
                w = w * wlen % R_MOD
        length <<= 1


def intt_inplace(a: List[int], omega: int) -> None:
    n = len(a)
    if n == 0:
        return
    ntt_inplace(a, fr_inv(int(omega) % R_MOD))
    inv_n = fr_inv(n)
    for i in range(n):
        a[i] = a[i] % R_MOD * inv_n % R_MOD


def evals_from_coeffs_on_roots(coeffs: Sequence[int], n: int, omega: int) -> List[int]:
    if n <= 0 or n & (n - 1):
        raise ValueError("n must be a power of two")
    out = [0] * n
    for i in range(min(n, len(coeffs))):
        out[i] = int(coeffs[i]) % R_MOD
    ntt_inplace(out, omega)
    return out


def coeffs_from_evals_on_roots(evals: Sequence[int], omega: int) -> List[int]:
    if len(evals) == 0 or len(evals) & (len(evals) - 1):
        raise ValueError("len(evals) must be a power of two")
    out = [int(x) % R_MOD for x in evals]
    intt_inplace(out, omega)
    return out


def evals_from_coeffs_on_coset(coeffs: Sequence[int], *, n: int, omega: int, shift: int) -> List[int]:
    if n <= 0 or n & (n - 1):
        raise ValueError("n must be a power of two")
    ss = int(shift) % R_MOD
    out = [0] * n
    pw = 1
    for i in range(min(n, len(coeffs))):
        out[i] = int(coeffs[i]) % R_MOD * pw % R_MOD
        pw = pw * ss % R_MOD
    ntt_inplace(out, omega)
    return out


def coeffs_from_evals_on_coset(evals: Sequence[int], *, omega: int, shift: int) -> List[int]:
    if len(evals) == 0 or len(evals) & (len(evals) - 1):
        raise ValueError("len(evals) must be a power of two")
    ss = int(shift) % R_MOD
    inv_s = fr_inv(ss) if ss != 0 else 0
    out = [int(x) % R_MOD for x in evals]
    intt_inplace(out, omega)
    pw = 1
    for i in range(len(out)):
        out[i] = out[i] * pw % R_MOD
        pw = pw * inv_s % R_MOD if inv_s != 0 else 0
    return out


def dft_naive(a: Sequence[int], omega: int) -> List[int]:
    """out[i] = sum_j a_j omega^(ij) — the definition."""
    n = len(a)
    w = int(omega) % R_MOD
    out = []
    for i in range(n):
        wi = pow(w, i, R_MOD)
        acc = 0
        x = 1
        for j in range(n):
            acc += int(a[j]) * x
            x = x * wi % R_MOD
        out.append(acc % R_MOD)
    return out


def poly_eval(coeffs: Sequence[int], x: int) -> int:
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + int(c)) % R_MOD
    return acc
