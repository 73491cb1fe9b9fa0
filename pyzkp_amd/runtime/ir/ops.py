"""Operator types of the runtime graph — the reference's OpType enum
(runtime/ir/ops.py:6-26), same members and string values, so graphs built
against the reference's API run unchanged."""
from __future__ import annotations

from enum import Enum


class OpType(str, Enum):
    TO_DEVICE = "to_device"
    FROM_DEVICE = "from_device"
    ROOTS_EVALS_FROM_COEFFS = "roots_evals_from_coeffs"
    ROOTS_COEFFS_FROM_EVALS = "roots_coeffs_from_evals"
    COSET_EVALS_FROM_COEFFS = "coset_evals_from_coeffs"
    COSET_COEFFS_FROM_EVALS = "coset_coeffs_from_evals"
    BATCH_INV = "batch_inv"
    POINTWISE_MUL = "pointwise_mul"
    POLY_MUL_NTT = "poly_mul_ntt"
    POLY_SUB = "poly_sub"
    DIV_XN_MINUS_1 = "div_xn_minus_1"
    PLONK_T_QUOTIENT_EVALS = "plonk_t_quotient_evals"
    MSM_G1 = "msm_g1"
    MSM_G2 = "msm_g2"
    MSM_G1_BATCH = "msm_g1_batch"
    KZG_COMMIT = "kzg_commit"
    KZG_OPEN = "kzg_open"
    KZG_OPEN_PREP_BATCH = "kzg_open_prep_batch"
    KZG_BATCH_COMMIT = "kzg_batch_commit"
    KZG_BATCH_OPEN = "kzg_batch_open"
