"""Host-side BN254 G2 for the operator layer (CPU threads in libzkhip).

Host G2 points follow the reference's py_ecc convention (crypto/ecc/bn254.py:7-54):
homogeneous (X, Y, Z) with Fq2 coordinates, here as (c0, c1) int pairs;
normalised points have Z = (1, 0); the identity is ((1,0), (1,0), (0,0)).
"""
from __future__ import annotations

import ctypes as ct
import threading
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._native import load_library
from .field import FQ_MODULUS, FR_MODULUS

Q = FQ_MODULUS
F2 = Tuple[int, int]
G2Point = Tuple[F2, F2, F2]
G2_ZERO: G2Point = ((1, 0), (1, 0), (0, 0))
G2_GENERATOR: G2Point = (
    (10857046999023057135944570762232829481370756359578518086990519993285655852781,
     11559732032986387107991004021392285783925812861821192530917403151452391805634),
    (8495653923123431417604973247489272438418190587263600148770280649306958101930,
     4082367875863433681332203403145435568316851327593401208105741076214120093531),
    (1, 0),
)


def _f2(v) -> F2:
    if hasattr(v, "coeffs"):  # py_ecc FQ2
        return (int(v.coeffs[0]) % Q, int(v.coeffs[1]) % Q)
    return (int(v[0]) % Q, int(v[1]) % Q)


def f2_mul(a: F2, b: F2) -> F2:
    return ((a[0] * b[0] - a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def f2_inv(a: F2) -> F2:
    t = pow((a[0] * a[0] + a[1] * a[1]) % Q, -1, Q)
    return (a[0] * t % Q, (-a[1]) * t % Q)


def normalize(p) -> Optional[Tuple[F2, F2]]:
    X, Y, Z = (_f2(c) for c in p)
    if Z == (0, 0):
        return None
    zi = f2_inv(Z)
    return (f2_mul(X, zi), f2_mul(Y, zi))


def g2_eq(p, q) -> bool:
    return normalize(p) == normalize(q)


def _pack(points) -> np.ndarray:
    out = np.zeros((len(points), 16), dtype="<u8")
    for i, p in enumerate(points):
        a = normalize(p)
        if a is None:
            continue
        vals = [a[0][0], a[0][1], a[1][0], a[1][1]]
        out[i] = np.frombuffer(b"".join(v.to_bytes(32, "little") for v in vals), dtype="<u8")
    return out


def _unpack(words: np.ndarray) -> G2Point:
    b = words.tobytes()
    v = [int.from_bytes(b[32 * k:32 * k + 32], "little") for k in range(4)]
    if not any(v):
        return G2_ZERO
    return ((v[0], v[1]), (v[2], v[3]), (1, 0))


def _scal(scalars) -> np.ndarray:
    return np.frombuffer(b"".join((int(s) % FR_MODULUS).to_bytes(32, "little") for s in scalars),
                         dtype="<u8").copy() if len(scalars) else np.zeros(4, dtype="<u8")


class G2Bases:
    """Packed affine host copy of a G2 point sequence (reused across MSMs).

    MSMs over it run on a prepared base set (zk_g2_bases_create: the points
    converted once for the IFMA path), made at the first MSM and kept until
    free() or garbage collection."""

    def __init__(self, points) -> None:
        self.words = _pack(points)
        self.n = len(points)
        self._prep = None
        self._lock = threading.Lock()

    def prepared(self, threads: int = 0):
        """The native prepared set (a handle), created on first use."""
        if self._prep is None:
            with self._lock:
                if self._prep is None:
                    h = ct.c_void_p()
                    w = np.ascontiguousarray(self.words, dtype="<u8")
                    rc = load_library().zk_g2_bases_create(w.ctypes.data, self.n, threads, ct.byref(h))
                    if rc:
                        raise RuntimeError(f"zk_g2_bases_create failed ({rc})")
                    self._prep = h
        return self._prep

    def free(self) -> None:
        with self._lock:
            if self._prep is not None and self._prep.value:
                load_library().zk_g2_bases_destroy(self._prep)
            self._prep = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _msm(bases: "G2Bases", s: np.ndarray, n: int, window_bits: int, threads: int, prepared: bool) -> G2Point:
    out = np.zeros(16, dtype="<u8")
    lib = load_library()
    if prepared:
        rc = lib.zk_g2_msm_prepared(bases.prepared(threads), s.ctypes.data, n, window_bits, threads, out.ctypes.data)
    else:
        rc = lib.zk_g2_msm_host(bases.words.ctypes.data, s.ctypes.data, n, window_bits, threads, out.ctypes.data)
    if rc:
        raise ValueError("g2 msm failed")
    return _unpack(out)


def g2_msm(points, scalars: Sequence[int], window_bits: int = 0, threads: int = 0) -> G2Point:
    """sum_i s_i Q_i (crypto/msm/pippenger.py:67-121 semantics).  A G2Bases
    runs on its prepared set; a point list is converted for this call."""
    keep = isinstance(points, G2Bases)
    bases = points if keep else G2Bases(points)
    n = len(scalars)
    if n > bases.n:
        raise ValueError("points and scalars length mismatch")
    return _msm(bases, _scal(scalars), n, window_bits, threads, prepared=keep)


def g2_msm_words(bases: "G2Bases", scalar_words: np.ndarray, n: int, window_bits: int = 0,
                 threads: int = 0, prepared: bool = True) -> G2Point:
    """Same, with canonical scalars as (n, 4) uint64 words; on the prepared
    set unless prepared=False."""
    if n > bases.n:
        raise ValueError("points and scalars length mismatch")
    s = np.ascontiguousarray(scalar_words, dtype="<u8")
    if s.ndim != 2 or s.shape[1] != 4 or s.shape[0] < n:
        raise ValueError(f"scalar words must have shape (>= {n}, 4), got {tuple(s.shape)}")
    return _msm(bases, s, n, window_bits, threads, prepared)


def g2_mul_many(q, scalars, threads: int = 0) -> np.ndarray:
    """[k_i] q for many k_i: (n, 16) packed affine words."""
    n = len(scalars)
    out = np.zeros((max(n, 1), 16), dtype="<u8")
    qw = _pack([q])
    s = _scal(scalars)
    rc = load_library().zk_g2_mul_fixed_host(qw.ctypes.data, s.ctypes.data, n, threads, out.ctypes.data)
    if rc:
        raise ValueError("g2 fixed-base multiplication failed")
    return out[:n]


def g2_mul_many_words(q, scalar_words: np.ndarray, threads: int = 0) -> np.ndarray:
    n = scalar_words.shape[0]
    out = np.zeros((max(n, 1), 16), dtype="<u8")
    qw = _pack([q])
    s = np.ascontiguousarray(scalar_words, dtype="<u8")
    rc = load_library().zk_g2_mul_fixed_host(qw.ctypes.data, s.ctypes.data, n, threads, out.ctypes.data)
    if rc:
        raise ValueError("g2 fixed-base multiplication failed")
    return out[:n]


def g2_mul(q, k: int) -> G2Point:
    return _unpack(g2_mul_many(q, [k])[0])


def g2_add(p, q) -> G2Point:
    out = np.zeros(16, dtype="<u8")
    a, b = _pack([p]), _pack([q])
    load_library().zk_g2_add_host(a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return _unpack(out)


def from_words(words: np.ndarray) -> G2Bases:
    b = G2Bases([])
    b.words = np.ascontiguousarray(words, dtype="<u8")
    b.n = words.shape[0]
    b._prep = None
    return b


def point_from_words(w: np.ndarray) -> G2Point:
    return _unpack(np.ascontiguousarray(w, dtype="<u8"))
