"""CPU oracle for the BN254 MSM / Fr NTT hot path — TEST INFRASTRUCTURE ONLY.

This package restates the reference's CPU algorithms so the HIP path can be
checked bit-for-bit.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import or run anything here, and
only as the checker / the timed CPU baseline — never as the thing measured
or shipped.  The product (``pyzkp_amd``) must not import it.

Contents (each function cites the reference file:line it follows):
  * ``field``  — Fr / Fq constants and helpers (crypto/field/fr.py,
                 runtime/kernels/cpu/kernels.py:46-50 Montgomery packing)
  * ``ntt``    — radix-2 DIT NTT / iNTT / coset forms (crypto/poly/ntt.py)
  * ``curve``  — G1 group law (crypto/ecc/bn254.py over py_ecc)
  * ``msm``    — naive MSM, unsigned-window Pippenger, signed digits, CSR
                 bucketing (crypto/msm/*.py, runtime/metal/msm.py:145-177,
                 runtime/metal/ecc_source.py:437-566)
  * ``gen``    — the deterministic input generator used by tests and bench
  * ``c/``     — a C restatement of the same NTT and Pippenger for large
                 sizes and for the CPU baseline timing (``liboracle.so``)

How each function is pinned (golden vectors, closed forms, brute force) is
listed in the module docstrings; every function has at least one such check
in ``tests/test_oracle.py``.
"""
