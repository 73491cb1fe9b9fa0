"""ctypes wrapper of the C oracle (oracle/_build/liboracle.so) — test
infrastructure: large-size parity and the timed CPU baseline.

Pinned by tests/test_oracle_c.py against the Python oracle (which is pinned
by golden vectors and closed forms) on sizes the Python oracle finishes
quickly, and by the closed form [sum s_i k_i] G for MSMs.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ct.CDLL(LIB)
        P = ct.c_void_p
        L.oracle_ntt.argtypes = [P, ct.c_uint64, P, ct.c_uint32, P, ct.c_int, P, ct.c_int]
        L.oracle_msm_pippenger.argtypes = [P, P, ct.c_uint64, ct.c_int, P, ct.POINTER(ct.c_int), ct.c_int]
        L.oracle_g1_mul.argtypes = [P, P, P, ct.POINTER(ct.c_int)]
        L.oracle_max_threads.argtypes = []
        L.oracle_poly_eval.argtypes = [P, ct.c_uint64, P, P]
        _lib = L
    return _lib


def _words(xs: Sequence[int]) -> np.ndarray:
    return np.frombuffer(b"".join(int(x).to_bytes(32, "little") for x in xs), dtype="<u8").copy()


def _ints(a: np.ndarray) -> List[int]:
    b = a.tobytes()
    return [int.from_bytes(b[i:i + 32], "little") for i in range(0, len(b), 32)]


def ntt_words(inp: np.ndarray, in_len: int, log_n: int, omega: int, inverse: bool,
              shift: Optional[int] = None, threads: int = 1) -> np.ndarray:
    """inp: canonical uint64 words (flat, 4 per element)."""
    n = 1 << log_n
    out = np.zeros(4 * n, dtype="<u8")
    inp = np.ascontiguousarray(inp, dtype="<u8")
    om = _words([omega])
    sh = _words([shift]) if shift is not None else None
    rc = lib().oracle_ntt(inp.ctypes.data, in_len, out.ctypes.data, log_n, om.ctypes.data, int(inverse),
                          sh.ctypes.data if sh is not None else None, threads)
    if rc:
        raise RuntimeError("oracle_ntt failed")
    return out


def ntt(vals: Sequence[int], n: int, omega: int, inverse: bool, shift: Optional[int] = None,
        threads: int = 1) -> List[int]:
    log_n = n.bit_length() - 1
    w = _words(vals) if len(vals) else np.zeros(4, dtype="<u8")
    return _ints(ntt_words(w, len(vals), log_n, omega, inverse, shift, threads))


def msm_pippenger_words(bases_xy: np.ndarray, scalars: np.ndarray, n: int, window_bits: int,
                        threads: int = 1) -> Optional[Tuple[int, int]]:
    """bases_xy: canonical affine words (8 per point, x = y = 0 identity);
    scalars: canonical words (4 per scalar)."""
    out = np.zeros(8, dtype="<u8")
    inf = ct.c_int(0)
    b = np.ascontiguousarray(bases_xy, dtype="<u8")
    s = np.ascontiguousarray(scalars, dtype="<u8")
    rc = lib().oracle_msm_pippenger(b.ctypes.data, s.ctypes.data, n, window_bits, out.ctypes.data,
                                    ct.byref(inf), threads)
    if rc:
        raise RuntimeError("oracle_msm_pippenger failed")
    if inf.value:
        return None
    v = _ints(out)
    return (v[0], v[1])


def msm_pippenger(points, scalars, window_bits: int, threads: int = 1):
    xy = []
    for p in points:
        xy += [0, 0] if p is None else [p[0], p[1]]
    return msm_pippenger_words(_words(xy), _words(scalars), len(points), window_bits, threads)


def g1_mul(p, k: int):
    out = np.zeros(8, dtype="<u8")
    inf = ct.c_int(0)
    xy = _words([0, 0] if p is None else [p[0], p[1]])
    kk = _words([k])
    lib().oracle_g1_mul(xy.ctypes.data, kk.ctypes.data, out.ctypes.data, ct.byref(inf))
    if inf.value:
        return None
    v = _ints(out)
    return (v[0], v[1])


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def poly_eval_words(coeffs: np.ndarray, n: int, x: int) -> int:
    out = np.zeros(4, dtype="<u8")
    c = np.ascontiguousarray(coeffs, dtype="<u8")
    xx = _words([x])
    lib().oracle_poly_eval(c.ctypes.data, n, xx.ctypes.data, out.ctypes.data)
    return _ints(out)[0]
