"""Groth16 verification (reference: protocols/groth16/verify.py:11-23):
e(A, B) == e(alpha, beta) e(vk_x, gamma) e(C, delta), checked as one
product of four pairings with a single final exponentiation."""
from __future__ import annotations

from typing import Sequence

from ..curve import G1_ZERO, g1_add, g1_mul_native as g1_mul, g1_neg
from ..field import FR_MODULUS
from ..pairing import pairing_check
from .types import Proof, VerifyingKey


def verify(vk: VerifyingKey, public_values: Sequence[int], proof: Proof) -> bool:
    if len(public_values) != len(vk.ic):
        return False
    vk_x = G1_ZERO
    for sc, p in zip(public_values, vk.ic):
        vk_x = g1_add(vk_x, g1_mul(p, int(sc) % FR_MODULUS))
    return pairing_check([(proof.a, proof.b), (g1_neg(vk.alpha_g1), vk.beta_g2),
                          (g1_neg(vk_x), vk.gamma_g2), (g1_neg(proof.c), vk.delta_g2)])
