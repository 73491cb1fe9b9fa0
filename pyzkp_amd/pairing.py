"""BN254 pairing on the host (libzkhip CPU code): the verifier's primitive
(crypto/pairing/bn254.py:10-11 in the reference, over py_ecc)."""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

from . import curve as C1
from . import g2 as C2
from ._native import load_library


def _g1_words(points) -> np.ndarray:
    out = np.zeros((max(len(points), 1), 8), dtype="<u8")
    for i, p in enumerate(points):
        a = C1.normalize(p)
        if a is None:
            continue
        out[i] = np.frombuffer(a[0].to_bytes(32, "little") + a[1].to_bytes(32, "little"), dtype="<u8")
    return out


def _g2_words(points) -> np.ndarray:
    w = C2._pack(points)
    return w if len(points) else np.zeros((1, 16), dtype="<u8")


def pairing_product(pairs: Sequence[Tuple[object, object]]) -> Tuple[int, ...]:
    """prod e(P_i, Q_i) as 12 canonical Fq coefficients."""
    g1 = _g1_words([p for p, _ in pairs])
    g2 = _g2_words([q for _, q in pairs])
    out = np.zeros(48, dtype="<u8")
    rc = load_library().zk_pairing_product(g1.ctypes.data, g2.ctypes.data, len(pairs), out.ctypes.data)
    if rc:
        raise ValueError("pairing failed")
    b = out.tobytes()
    return tuple(int.from_bytes(b[32 * k:32 * k + 32], "little") for k in range(12))


def pairing(p, q) -> Tuple[int, ...]:
    return pairing_product([(p, q)])


def pairing_check(pairs: Sequence[Tuple[object, object]]) -> bool:
    """True iff prod e(P_i, Q_i) == 1."""
    import ctypes as ct
    g1 = _g1_words([p for p, _ in pairs])
    g2 = _g2_words([q for _, q in pairs])
    ok = ct.c_int(0)
    rc = load_library().zk_pairing_check(g1.ctypes.data, g2.ctypes.data, len(pairs), ct.byref(ok))
    if rc:
        raise ValueError("pairing failed")
    return bool(ok.value)
