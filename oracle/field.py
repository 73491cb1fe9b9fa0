"""BN254 field constants and Montgomery packing (oracle; test infrastructure).

FR_MODULUS is py_ecc's ``curve_order`` (crypto/field/fr.py:5-7); the
Montgomery radix R = 2^256 and the 4 x u64 little-endian packing follow
runtime/kernels/cpu/kernels.py:46-50, 135-140, 168-170.

Pinned by: closed-form checks in tests/test_oracle.py (r and q primality
witnesses via Fermat, 2-adicity 28 of r - 1, 5 a non-residue mod r,
ROOT_2_28 = 5^((r-1)/2^28) of order exactly 2^28).
"""
from __future__ import annotations

R_MOD = 21888242871839275222246405745257275088548364400416034343698204186575808495617
Q_MOD = 21888242871839275222246405745257275088696311157297823662689037894645226208583
MONT_R = 1 << 256
FR_GENERATOR = 5  # multiplicative generator used as the coset shift (crypto/poly/ntt.py:8 comment, tests shift 5)
# crypto/poly/ntt.py:8
ROOT_2_28 = 19103219067921713944291392827692070036145651957329286315305642004821462161904
MAX_ORDER_LOG = 28


def to_mont(x: int, p: int) -> int:
    return (int(x) % p) * MONT_R % p


def from_mont(x: int, p: int) -> int:
    return int(x) * pow(MONT_R, -1, p) % p


def to_words(x: int) -> list[int]:
    return [(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)]


def from_words(w) -> int:
    return int(w[0]) | (int(w[1]) << 64) | (int(w[2]) << 128) | (int(w[3]) << 192)


def fr_inv(a: int) -> int:
    """crypto/field/fr.py:33-37 (raises on 0)."""
    aa = a % R_MOD
    if aa == 0:
        raise ZeroDivisionError
    return pow(aa, R_MOD - 2, R_MOD)
