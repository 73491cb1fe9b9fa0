"""Helpers for the GPU parity tests (device upload / download of Fr, G1)."""
from __future__ import annotations

from typing import List, Sequence

from pyzkp_amd.field import (FQ_MODULUS, FR_MODULUS, fq_to_mont, fr_to_mont, pack_ints, unpack_ints,
                             fr_from_mont)


def upload_fr_mont(ctx, xs: Sequence[int]):
    """Canonical ints -> device Montgomery buffer (converted on the device)."""
    n = len(xs)
    buf = ctx.alloc(max(32 * n, 32))
    if n:
        ctx.h2d(buf.ptr, pack_ints(int(x) % FR_MODULUS for x in xs), 32 * n)
        ctx.fr_to_mont(buf.ptr, buf.ptr, n)
        ctx.sync()
    return buf


def download_fr(ctx, ptr: int, n: int) -> List[int]:
    """Device Montgomery buffer -> canonical ints (converted on the host)."""
    raw = ctx.download(ptr, 32 * n)
    return [fr_from_mont(v) for v in unpack_ints(raw)]


def upload_bases(ctx, pts):
    """Affine (x, y) tuples or None -> 64-byte affine Montgomery device buffer."""
    b = bytearray()
    for p in pts:
        if p is None:
            b += bytes(64)
        else:
            b += fq_to_mont(p[0]).to_bytes(32, "little") + fq_to_mont(p[1]).to_bytes(32, "little")
    return ctx.upload(bytes(b))


def msm_affine(ctx, bases_buf, scalars_buf, n, c=0):
    from pyzkp_amd.curve import from_jacobian_mont

    X, Y, Z = ctx.msm_g1(bases_buf.ptr, scalars_buf.ptr, n, c)
    p = from_jacobian_mont(X, Y, Z)
    return None if p[2] == 0 else (p[0], p[1])
