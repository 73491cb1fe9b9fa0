"""Host-side BN254 constants and 32-byte little-endian packing.

FR_MODULUS / FQ_MODULUS are the BN254 scalar and base field moduli
(crypto/field/fr.py:7 takes r from py_ecc); Montgomery radix R = 2^256 and
the 4 x u64 little-endian layout follow runtime/kernels/cpu/kernels.py:46-50.
"""
from __future__ import annotations

from typing import Iterable, List

import numpy as np

FR_MODULUS = 21888242871839275222246405745257275088548364400416034343698204186575808495617
FQ_MODULUS = 21888242871839275222246405745257275088696311157297823662689037894645226208583
MONT_R = 1 << 256
# 5^((r-1)/2^28), a primitive 2^28-th root of unity (crypto/poly/ntt.py:8)
ROOT_2_28 = 19103219067921713944291392827692070036145651957329286315305642004821462161904
MAX_ORDER_LOG = 28
FR_GENERATOR = 5

_FR_RINV = pow(MONT_R, -1, FR_MODULUS)
_FQ_RINV = pow(MONT_R, -1, FQ_MODULUS)


def fr_to_mont(x: int) -> int:
    return int(x) % FR_MODULUS * MONT_R % FR_MODULUS


def fr_from_mont(x: int) -> int:
    return int(x) * _FR_RINV % FR_MODULUS


def fq_to_mont(x: int) -> int:
    return int(x) % FQ_MODULUS * MONT_R % FQ_MODULUS


def fq_from_mont(x: int) -> int:
    return int(x) * _FQ_RINV % FQ_MODULUS


def fr_inv(a: int) -> int:
    aa = int(a) % FR_MODULUS
    if aa == 0:
        raise ZeroDivisionError
    return pow(aa, FR_MODULUS - 2, FR_MODULUS)


def omega_for_size(n: int) -> int:
    """A primitive n-th root of unity: ROOT_2_28^(2^(28 - log2 n))
    (crypto/poly/ntt.py:13-24)."""
    if n <= 0 or n & (n - 1):
        raise ValueError("n must be a power of two")
    logn = n.bit_length() - 1
    if logn > MAX_ORDER_LOG:
        raise ValueError("domain too large for BN254 root of unity")
    return pow(ROOT_2_28, 1 << (MAX_ORDER_LOG - logn), FR_MODULUS)


def pack_ints(xs: Iterable[int]) -> bytes:
    """Little-endian 32-byte words (values must already be < 2^256)."""
    return b"".join(int(x).to_bytes(32, "little") for x in xs)


def unpack_ints(b: bytes) -> List[int]:
    mv = memoryview(b)
    return [int.from_bytes(mv[i:i + 32], "little") for i in range(0, len(b), 32)]


def words_to_ints(a: np.ndarray) -> List[int]:
    """(n, 4) uint64 little-endian words -> Python ints."""
    return unpack_ints(np.ascontiguousarray(a, dtype="<u8").tobytes())


def ints_to_words(xs: Iterable[int]) -> np.ndarray:
    return np.frombuffer(pack_ints(xs), dtype="<u8").reshape(-1, 4).copy()
