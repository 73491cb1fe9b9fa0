"""PLONK quotient on the device (reference: protocols/plonk/prove.py:300-424).

The same graph as the reference's quotient step, run by this package's
Executor with the hip kernels: 15 coset NTTs of size m = 4n
(COSET_EVALS_FROM_COEFFS; the coefficient vectors are zero-padded on the
device), the fused PLONK_T_QUOTIENT_EVALS numerator, BATCH_INV of Z_H,
POINTWISE_MUL and the coset iNTT, with one FROM_DEVICE at the end instead of
a round trip per operator.  The shift is an argument (`draw_shift` follows
the reference's draw, prove.py:323-327); L1's coefficients are the inverse
NTT of [1, 0, ..., 0], i.e. n^-1 in every position.
"""
from __future__ import annotations

import secrets
from typing import Dict, Optional, Sequence, Tuple

from ..field import FR_MODULUS, fr_inv, omega_for_size
from ..runtime import Backend, Device, DType, Executor, Graph, KernelRegistry, OpType, RuntimeConfig
from ..runtime.kernels.hip import register_hip_kernels

_INPUTS = ("a", "b", "c", "z", "pi", "ql", "qr", "qm", "qo", "qc", "s1", "s2", "s3")


def draw_shift(n: int) -> int:
    """A coset shift whose n-th power avoids {1, u, u^2, u^3}, u = omega_4n^n
    (prove.py:323-327), so Z_H never vanishes on the extended coset."""
    u = pow(omega_for_size(4 * n), n, FR_MODULUS)
    bad = {1, u, u * u % FR_MODULUS, u * u % FR_MODULUS * u % FR_MODULUS}
    while True:
        shift = secrets.randbelow(FR_MODULUS - 1) + 1
        if pow(shift, n, FR_MODULUS) not in bad:
            return shift


def compute_quotient(coeffs: Dict[str, Sequence[int]], *, n: int, omega: int, k1: int, k2: int, alpha: int,
                     beta: int, gamma: int, shift: Optional[int] = None, context=None, pool=None,
                     trace=None, check_degree: bool = True):
    """t1, t2, t3 (n coefficients each) for the wire / permutation /
    selector coefficient vectors in `coeffs` (keys a, b, c, z, pi, ql, qr,
    qm, qo, qc, s1, s2, s3; host ints).  Raises ValueError("quotient degree
    too large") as the reference does; with check_degree=False returns all
    4n coefficients instead."""
    m = 4 * n
    omega_m = omega_for_size(m)
    if shift is None:
        shift = draw_shift(n)
    shift = int(shift) % FR_MODULUS
    reg = KernelRegistry()
    register_hip_kernels(reg)
    exe = Executor(registry=reg)
    g = Graph()
    for k in _INPUTS:
        g.add_buffer(id=f"{k}_coeff", device=Device.CPU, dtype=DType.FR, data=list(coeffs[k]))
    g.add_buffer(id="l1_coeff", device=Device.CPU, dtype=DType.FR, data=[fr_inv(n)] * n)
    coset = {"n": m, "omega": omega_m, "shift": shift}
    for k in _INPUTS:
        g.add_node(op=OpType.COSET_EVALS_FROM_COEFFS, inputs=[f"{k}_coeff"], outputs=[f"{k}_ext"], attrs=coset)
    g.add_node(op=OpType.COSET_EVALS_FROM_COEFFS, inputs=["z_coeff"], outputs=["zshift_ext"],
               attrs={"n": m, "omega": omega_m, "shift": shift * int(omega) % FR_MODULUS})
    g.add_node(op=OpType.COSET_EVALS_FROM_COEFFS, inputs=["l1_coeff"], outputs=["l1_ext"], attrs=coset)
    q_attrs = {"n": n, "m": m, "omega_m": omega_m, "shift": shift, "shift_n": pow(shift, n, FR_MODULUS),
               "alpha": int(alpha) % FR_MODULUS, "beta": int(beta) % FR_MODULUS,
               "gamma": int(gamma) % FR_MODULUS, "k1": int(k1) % FR_MODULUS, "k2": int(k2) % FR_MODULUS}
    g.add_node(op=OpType.PLONK_T_QUOTIENT_EVALS,
               inputs=[f"{k}_ext" for k in _INPUTS] + ["zshift_ext", "l1_ext"],
               outputs=["num_ext", "zh_ext"], attrs=q_attrs)
    g.add_node(op=OpType.BATCH_INV, inputs=["zh_ext"], outputs=["inv_zh_ext"])
    g.add_node(op=OpType.POINTWISE_MUL, inputs=["num_ext", "inv_zh_ext"], outputs=["t_ext"])
    g.add_node(op=OpType.COSET_COEFFS_FROM_EVALS, inputs=["t_ext"], outputs=["t_coeff_full_m"],
               attrs={"omega": omega_m, "shift": shift})
    g.add_node(op=OpType.FROM_DEVICE, inputs=["t_coeff_full_m"], outputs=["t_coeff_full"])
    cfg = RuntimeConfig(backend=Backend.HIP)
    exe.run(g, pool=pool, trace=trace, keep=["t_coeff_full"], context=context, runtime_config=cfg)
    t_full = g.buffers["t_coeff_full"].data
    if not check_degree:
        return t_full
    for v in t_full[3 * n:]:
        if v % FR_MODULUS != 0:
            raise ValueError("quotient degree too large")
    t = list(t_full[:3 * n]) + [0] * max(0, 3 * n - len(t_full))
    return tuple(t[0:n]), tuple(t[n:2 * n]), tuple(t[2 * n:3 * n])
