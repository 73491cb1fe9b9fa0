"""HIP backend operators — the drop-in replacement of the reference's Metal
kernels (runtime/kernels/metal/kernels.py) behind the same KernelFn
contract (runtime/kernels/registry.py:17):

  ctx = {"graph", "node", "inputs": [Buffer], "attrs", "pool", "context"}
  return {"outputs": {buffer_id: Buffer}}

Registered under (op, Device.HIP, Backend.HIP):
  ROOTS_EVALS_FROM_COEFFS  {n, omega}         metal kernels.py:500-598 / cpu kernels.py:175-190
  ROOTS_COEFFS_FROM_EVALS  {omega}            metal kernels.py:704-827 / cpu kernels.py:193-206
  COSET_EVALS_FROM_COEFFS  {n, omega, shift}  metal kernels.py:225-303 / cpu kernels.py:209-228
  COSET_COEFFS_FROM_EVALS  {omega, shift}     metal kernels.py:306-381 / cpu kernels.py:231-250
  POINTWISE_MUL, POLY_SUB                     metal kernels.py:122-222
  MSM_G1                                      metal kernels.py:84-120
  KZG_BATCH_COMMIT, KZG_COMMIT, MSM_G1_BATCH  metal kernels.py:22-82, cpu kernels.py:338-496
  BATCH_INV                                   cpu kernels.py:253-269
  PLONK_T_QUOTIENT_EVALS                      cpu kernels.py:607-694
  FROM_DEVICE (device side)                   cpu kernels.py:75-77 registers it on METAL too
and, under (op, Device.CPU, Backend.HIP), the transfer ops TO_DEVICE /
FROM_DEVICE (cpu kernels.py:98-172), which pack on the host only what has
to cross PCIe and do the Montgomery conversion on the device.

Device FR buffers are HipBuffer(dtype="fr_mont_u64x4", n, dbuf) — the
layout of the reference's MetalBuffer (runtime/metal/buffer.py:8-15).
Errors follow the reference: ValueError for dtype / length / size
problems, RuntimeError when the HIP context is missing.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass
from typing import Any, Dict, List, Optional

from ...curve import G1_ZERO, affine_mont_bytes, from_jacobian_mont
from ...field import FR_MODULUS, fr_from_mont, fr_to_mont, unpack_ints
from ..ir.ops import OpType
from ..ir.types import Backend, Buffer, Device, DType
from .registry import KernelRegistry


@dataclass
class HipBuffer:
    dtype: str
    n: int
    dbuf: Any  # pyzkp_amd.DeviceBuffer
    host_cache: Optional[list] = None

    def __len__(self) -> int:
        return int(self.n)

    @property
    def ptr(self) -> int:
        return self.dbuf.ptr


@dataclass
class DeviceBases:
    """G1 bases resident in HBM as 64-byte affine Montgomery points, plus
    their prepared form (converted once, made on the first MSM)."""
    dbuf: Any
    n: int
    prepared: Any = None

    def __len__(self) -> int:
        return int(self.n)


def register_hip_kernels(registry: KernelRegistry) -> None:
    H, B = Device.HIP, Backend.HIP
    registry.register(OpType.TO_DEVICE, Device.CPU, _to_device, backend=B)
    registry.register(OpType.FROM_DEVICE, Device.CPU, _from_device, backend=B)
    registry.register(OpType.FROM_DEVICE, H, _from_device, backend=B)
    registry.register(OpType.ROOTS_EVALS_FROM_COEFFS, H, _roots_evals_from_coeffs, backend=B)
    registry.register(OpType.ROOTS_COEFFS_FROM_EVALS, H, _roots_coeffs_from_evals, backend=B)
    registry.register(OpType.COSET_EVALS_FROM_COEFFS, H, _coset_evals_from_coeffs, backend=B)
    registry.register(OpType.COSET_COEFFS_FROM_EVALS, H, _coset_coeffs_from_evals, backend=B)
    registry.register(OpType.POINTWISE_MUL, H, _pointwise_mul, backend=B)
    registry.register(OpType.POLY_SUB, H, _poly_sub, backend=B)
    registry.register(OpType.MSM_G1, H, _msm_g1, backend=B)
    registry.register(OpType.MSM_G1_BATCH, H, _msm_g1_batch, backend=B)
    registry.register(OpType.MSM_G2, H, _msm_g2, backend=B)
    registry.register(OpType.KZG_COMMIT, H, _kzg_commit, backend=B)
    registry.register(OpType.KZG_BATCH_COMMIT, H, _kzg_batch_commit, backend=B)
    registry.register(OpType.BATCH_INV, H, _batch_inv, backend=B)
    registry.register(OpType.PLONK_T_QUOTIENT_EVALS, H, _plonk_t_quotient_evals, backend=B)


# ------------------------------------------------------------------ helpers
def _hip(ctx: Dict[str, Any]):
    c = ctx.get("context")
    if c is None or getattr(c, "hip", None) is None:
        raise RuntimeError("hip kernels require a HipContext with a HIP device")
    return c.hip


def _alloc(ctx: Dict[str, Any], nbytes: int):
    hip = _hip(ctx)
    pool = ctx.get("pool")
    if pool is not None and hasattr(pool, "alloc_device"):
        return pool.alloc_device(hip, max(int(nbytes), 32))
    return hip.alloc(max(int(nbytes), 32))


def _stage_copy_in(ctx, dst_ptr: int, data: bytes) -> None:
    hip = _hip(ctx)
    if not data:
        return
    pool = ctx.get("pool")
    if pool is not None and hasattr(pool, "staging"):
        sp = pool.staging(hip, len(data))
        ct.memmove(sp, data, len(data))
        hip.h2d(dst_ptr, sp, len(data))
    else:
        hip.h2d(dst_ptr, data, len(data))
    hip.sync()


def upload_fr(ctx: Dict[str, Any], values) -> HipBuffer:
    """Canonical ints (any size, reduced mod r) -> device Montgomery buffer."""
    hip = _hip(ctx)
    n = len(values)
    dbuf = _alloc(ctx, 32 * n)
    if n:
        data = b"".join((int(x) % FR_MODULUS).to_bytes(32, "little") for x in values)
        _stage_copy_in(ctx, dbuf.ptr, data)
        hip.fr_to_mont(dbuf.ptr, dbuf.ptr, n)
    return HipBuffer(dtype="fr_mont_u64x4", n=n, dbuf=dbuf)


def download_fr(ctx: Dict[str, Any], hb: HipBuffer) -> List[int]:
    hip = _hip(ctx)
    n = int(hb.n)
    if n == 0:
        return []
    tmp = _alloc(ctx, 32 * n)
    hip.fr_from_mont(hb.ptr, tmp.ptr, n)
    pool = ctx.get("pool")
    if pool is not None and hasattr(pool, "staging"):
        sp = pool.staging(hip, 32 * n)
        hip.d2h(sp, tmp.ptr, 32 * n)
        hip.sync()
        raw = ct.string_at(sp, 32 * n)
        pool.release_device(tmp)
    else:
        raw = hip.download(tmp.ptr, 32 * n)
        tmp.free()
    return unpack_ints(raw)


def _fr_buffer(ctx, buf: Buffer, what: str) -> HipBuffer:
    if buf.dtype != DType.FR:
        raise ValueError(f"hip {what} supports FR only")
    if not isinstance(buf.data, HipBuffer):
        raise ValueError(f"hip {what} expects a HipBuffer")
    return buf.data


def _new_fr(ctx, n: int) -> HipBuffer:
    return HipBuffer(dtype="fr_mont_u64x4", n=int(n), dbuf=_alloc(ctx, 32 * int(n)))


def _pow2_log(n: int) -> int:
    if n <= 0 or n & (n - 1):
        raise ValueError("n must be a power of two")
    return n.bit_length() - 1


# ---------------------------------------------------------------- transfers
def _to_device(ctx: Dict[str, Any]) -> Dict[str, Any]:
    node = ctx["node"]
    inp: Buffer = ctx["inputs"][0]
    out_id = node.outputs[0]
    if inp.dtype != DType.FR:
        if inp.dtype in (DType.OBJ, DType.G1, DType.G2):
            # points stay host objects; MSM kernels keep device copies in their cache
            return {"outputs": {out_id: Buffer(id=out_id, device=Device.HIP, dtype=inp.dtype, data=inp.data)}}
        raise ValueError(f"to_device supports FR only, got {inp.dtype}")
    if isinstance(inp.data, HipBuffer):
        hb = inp.data
    else:
        hb = upload_fr(ctx, inp.data)
    return {"outputs": {out_id: Buffer(id=out_id, device=Device.HIP, dtype=DType.FR, data=hb,
                                       meta={"n": hb.n})}}


def _from_device(ctx: Dict[str, Any]) -> Dict[str, Any]:
    node = ctx["node"]
    inp: Buffer = ctx["inputs"][0]
    out_id = node.outputs[0]
    if inp.device == Device.CPU and not isinstance(inp.data, HipBuffer):
        return {"outputs": {out_id: Buffer(id=out_id, device=Device.CPU, dtype=inp.dtype, data=inp.data)}}
    if inp.dtype != DType.FR:
        if inp.dtype in (DType.OBJ, DType.G1, DType.G2):
            return {"outputs": {out_id: Buffer(id=out_id, device=Device.CPU, dtype=inp.dtype, data=inp.data)}}
        raise ValueError("from_device supports FR only")
    if not isinstance(inp.data, HipBuffer):
        raise ValueError("from_device expects a HipBuffer")
    vals = download_fr(ctx, inp.data)
    return {"outputs": {out_id: Buffer(id=out_id, device=Device.CPU, dtype=DType.FR, data=vals,
                                       meta={"n": len(vals)})}}


# ---------------------------------------------------------------------- NTT
def _ntt(ctx, inp: HipBuffer, n: int, omega: int, inverse: bool, shift: Optional[int], what: str) -> HipBuffer:
    hip = _hip(ctx)
    log_n = _pow2_log(n)
    out = _new_fr(ctx, n)
    hip.ntt(inp.ptr, inp.n, out.ptr, log_n, inverse, omega_mont=fr_to_mont(int(omega) % FR_MODULUS),
            shift_mont=None if shift is None else fr_to_mont(int(shift) % FR_MODULUS))
    return out


def _roots_evals_from_coeffs(ctx):
    node, inp = ctx["node"], ctx["inputs"][0]
    hb = _fr_buffer(ctx, inp, "roots_evals_from_coeffs")
    n = int(ctx["attrs"]["n"])
    omega = int(ctx["attrs"]["omega"])
    out = _ntt(ctx, hb, n, omega, False, None, "roots_evals_from_coeffs")
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.HIP, dtype=DType.FR, data=out,
                                    meta={"n": n, "omega": omega})}}


def _roots_coeffs_from_evals(ctx):
    node, inp = ctx["node"], ctx["inputs"][0]
    hb = _fr_buffer(ctx, inp, "roots_coeffs_from_evals")
    omega = int(ctx["attrs"]["omega"])
    out = _ntt(ctx, hb, int(hb.n), omega, True, None, "roots_coeffs_from_evals")
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.HIP, dtype=DType.FR, data=out,
                                    meta={"n": int(hb.n), "omega": omega})}}


def _coset_evals_from_coeffs(ctx):
    node, inp = ctx["node"], ctx["inputs"][0]
    hb = _fr_buffer(ctx, inp, "coset_evals_from_coeffs")
    n = int(ctx["attrs"]["n"])
    omega = int(ctx["attrs"]["omega"])
    shift = int(ctx["attrs"]["shift"])
    out = _ntt(ctx, hb, n, omega, False, shift, "coset_evals_from_coeffs")
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.HIP, dtype=DType.FR, data=out,
                                    meta={"n": n, "omega": omega, "shift": shift})}}


def _coset_coeffs_from_evals(ctx):
    node, inp = ctx["node"], ctx["inputs"][0]
    hb = _fr_buffer(ctx, inp, "coset_coeffs_from_evals")
    omega = int(ctx["attrs"]["omega"])
    shift = int(ctx["attrs"]["shift"])
    out = _ntt(ctx, hb, int(hb.n), omega, True, shift, "coset_coeffs_from_evals")
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.HIP, dtype=DType.FR, data=out,
                                    meta={"n": int(hb.n), "omega": omega, "shift": shift})}}


# ------------------------------------------------------------- element-wise
def _binary(ctx, what: str, fn_name: str):
    node = ctx["node"]
    a = _fr_buffer(ctx, ctx["inputs"][0], what)
    b = _fr_buffer(ctx, ctx["inputs"][1], what)
    if len(a) != len(b):
        raise ValueError("length mismatch")
    out = _new_fr(ctx, a.n)
    getattr(_hip(ctx), fn_name)(a.ptr, b.ptr, out.ptr, a.n)
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.HIP, dtype=DType.FR, data=out, meta={"n": a.n})}}


def _pointwise_mul(ctx):
    return _binary(ctx, "pointwise_mul", "fr_mul")


def _poly_sub(ctx):
    return _binary(ctx, "poly_sub", "fr_sub")


def _batch_inv(ctx):
    """BATCH_INV (cpu kernels.py:253-269, crypto/field/batch.py:8-38): element-wise
    inverse, 0 -> 0, by simultaneous inversion on the device."""
    node = ctx["node"]
    a = _fr_buffer(ctx, ctx["inputs"][0], "batch_inv")
    out = _new_fr(ctx, a.n)
    if a.n:
        _hip(ctx).fr_inv(a.ptr, out.ptr, a.n)
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.HIP, dtype=DType.FR, data=out, meta={"n": a.n})}}


_PLONK_Q_INPUTS = ("a", "b", "c", "z", "pi", "ql", "qr", "qm", "qo", "qc", "s1", "s2", "s3", "zshift", "l1")


def _plonk_t_quotient_evals(ctx):
    """PLONK_T_QUOTIENT_EVALS (cpu kernels.py:607-694): num and Z_H over the
    size-m coset in one fused device pass.  attrs: n, m, omega_m, shift,
    shift_n, alpha, beta, gamma, k1, k2 (canonical ints)."""
    node, attrs = ctx["node"], ctx["attrs"]
    n, m = int(attrs["n"]), int(attrs["m"])
    ins = [_fr_buffer(ctx, b, "plonk_t_quotient_evals") for b in ctx["inputs"][:15]]
    if len(ins) != 15:
        raise ValueError("plonk_t_quotient_evals needs 15 inputs")
    if not all(len(v) == m for v in ins):
        raise ValueError("eval length mismatch")
    mont = [fr_to_mont(int(attrs[k]) % FR_MODULUS)
            for k in ("omega_m", "shift", "shift_n", "alpha", "beta", "gamma", "k1", "k2")]
    num, zh = _new_fr(ctx, m), _new_fr(ctx, m)
    _hip(ctx).plonk_quotient_evals([b.ptr for b in ins], n, m, *mont, num.ptr, zh.ptr)
    o_num, o_zh = node.outputs[0], node.outputs[1]
    return {"outputs": {
        o_num: Buffer(id=o_num, device=Device.HIP, dtype=DType.FR, data=num, meta={"m": m}),
        o_zh: Buffer(id=o_zh, device=Device.HIP, dtype=DType.FR, data=zh, meta={"m": m}),
    }}


# ---------------------------------------------------------------------- MSM
# Device copies of host point sequences, keyed by the identity of the
# sequence object (the reference keys its fixed-base and SRS-prefix caches
# the same way: crypto/msm/pippenger.py:225-248, runtime/warmup.py:17-26,
# runtime/kernels/cpu/kernels.py:58-67).  The entry holds the object so its
# id cannot be reused while cached.
_BASES_CACHE: Dict[tuple, tuple] = {}


def device_bases(ctx, points, n_needed: int) -> DeviceBases:
    if isinstance(points, DeviceBases):
        if n_needed > points.n:
            raise ValueError("scalars length exceeds points length")
        return points
    hip = _hip(ctx)
    key = (id(hip), id(points))
    hit = _BASES_CACHE.get(key)
    if hit is not None and hit[0] is points and hit[1].n == len(points):
        db = hit[1]
    else:
        data = affine_mont_bytes(points)
        dbuf = hip.alloc(max(len(data), 64))
        if data:
            hip.h2d(dbuf.ptr, data, len(data))
            hip.sync()
        db = DeviceBases(dbuf=dbuf, n=len(points))
        _BASES_CACHE[key] = (points, db)
    if n_needed > db.n:
        raise ValueError("scalars length exceeds points length")
    return db


def _window_bits(ctx) -> int:
    attrs = ctx.get("attrs") or {}
    if "hip_window_bits" in attrs:
        return int(attrs["hip_window_bits"])
    cfg = getattr(ctx.get("context"), "config", None)
    return int(getattr(cfg, "hip_msm_window_bits", 0) or 0)


def msm_device(ctx, bases: DeviceBases, scalars: HipBuffer, n: int):
    hip = _hip(ctx)
    if bases.prepared is None or bases.prepared.ctx is not hip or not bases.prepared.handle:
        bases.prepared = hip.msm_bases_create(bases.dbuf.ptr, bases.n)
    X, Y, Z = hip.msm_g1_prepared(bases.prepared, scalars.ptr, n, _window_bits(ctx))
    return from_jacobian_mont(X, Y, Z)


def _msm_g1(ctx):
    node = ctx["node"]
    points: Buffer = ctx["inputs"][0]
    scalars: Buffer = ctx["inputs"][1]
    if scalars.dtype != DType.FR:
        raise ValueError("hip msm_g1 scalars must be FR")
    if not isinstance(scalars.data, HipBuffer):
        raise ValueError("hip msm_g1 expects scalars to be a HipBuffer")
    n = int(scalars.data.n)
    bases = device_bases(ctx, points.data, n)
    acc = msm_device(ctx, bases, scalars.data, n) if n else G1_ZERO
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.CPU, dtype=DType.G1, data=acc)}}


# Fixed-base sets (W tables 2^(c w) B_i of an SRS prefix) cached like the
# device bases, for nodes whose attrs ask for fixed-base MSMs (the
# reference's fixed_base / warmup_fixed_base attrs, runtime/warmup.py:84-150,
# crypto/msm/pippenger.py:217-342).  With them a batch of commitments over
# one SRS is one device launch.
_FIXED_CACHE: Dict[tuple, tuple] = {}


def fixed_set(ctx, points, n: int, window_bits: int = 0):
    hip = _hip(ctx)
    key = (id(points), int(n), int(window_bits), id(hip))
    ent = _FIXED_CACHE.get(key)
    if ent is None:
        bases = device_bases(ctx, points, n)
        ent = (points, hip.fixed_bases_create(bases.dbuf.ptr, n, window_bits))
        _FIXED_CACHE[key] = ent
    return ent[1]


def _commit_fixed(ctx, points, scalar_lists, max_len: int) -> list:
    attrs = ctx.get("attrs") or {}
    fb = fixed_set(ctx, points, max_len, int(attrs.get("fixed_base_window_bits", 0) or 0))
    sizes = [len(sc) for sc in scalar_lists]
    offs, acc = [], 0
    for m in sizes:
        offs.append(acc)
        acc += m
    hb = upload_fr(ctx, [v for sc in scalar_lists for v in sc])
    res = _hip(ctx).msm_g1_fixed_batch(fb, hb.ptr, offs, sizes)
    pool = ctx.get("pool")
    if pool is not None and hasattr(pool, "release_device"):
        pool.release_device(hb.dbuf)
    else:
        hb.dbuf.free()
    return [from_jacobian_mont(*r) if m else G1_ZERO for r, m in zip(res, sizes)]


def _commit_many(ctx, points, scalar_lists) -> list:
    """One MSM per scalar list over the prefix of `points` (host ints); with
    attrs["fixed_base"] all of them in one fixed-base batch launch."""
    max_len = max((len(s) for s in scalar_lists), default=0)
    if max_len > len(points):
        raise ValueError("SRS too small for polynomial degree")
    if max_len == 0:
        return [G1_ZERO] * len(scalar_lists)
    if (ctx.get("attrs") or {}).get("fixed_base"):
        return _commit_fixed(ctx, points, scalar_lists, max_len)
    bases = device_bases(ctx, points, max_len)
    outs = []
    for sc in scalar_lists:
        if len(sc) == 0:
            outs.append(G1_ZERO)
            continue
        hb = upload_fr(ctx, sc)
        outs.append(msm_device(ctx, bases, hb, len(sc)))
        pool = ctx.get("pool")
        if pool is not None and hasattr(pool, "release_device"):
            pool.release_device(hb.dbuf)
        else:
            hb.dbuf.free()
    return outs


def _msm_g1_batch(ctx):
    node = ctx["node"]
    points: Buffer = ctx["inputs"][0]
    lists: Buffer = ctx["inputs"][1]
    outs = _commit_many(ctx, points.data, [list(s) for s in lists.data])
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.CPU, dtype=DType.OBJ, data=outs)}}


def _srs_points(srs_obj):
    return srs_obj.g1_powers if hasattr(srs_obj, "g1_powers") else srs_obj


def _kzg_commit(ctx):
    node = ctx["node"]
    srs: Buffer = ctx["inputs"][0]
    coeffs: Buffer = ctx["inputs"][1]
    data = coeffs.data
    if isinstance(data, HipBuffer):
        pts = _srs_points(srs.data)
        if data.n > len(pts):
            raise ValueError("SRS too small for polynomial degree")
        cm = msm_device(ctx, device_bases(ctx, pts, data.n), data, data.n) if data.n else G1_ZERO
    else:
        cm = _commit_many(ctx, _srs_points(srs.data), [list(data)])[0]
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.CPU, dtype=DType.G1, data=cm)}}


def _kzg_batch_commit(ctx):
    node = ctx["node"]
    srs: Buffer = ctx["inputs"][0]
    polys: Buffer = ctx["inputs"][1]
    outs = _commit_many(ctx, _srs_points(srs.data), [list(p) for p in polys.data])
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.CPU, dtype=DType.OBJ, data=outs)}}


# ---------------------------------------------------------------------- MSM G2
# Device G2 base sets, cached by the identity of the host sequence like the
# G1 bases above.  The reference registers MSM_G2 on the CPU only
# (runtime/kernels/cpu/kernels.py:89, :338-348); here it runs on the device.
_G2_CACHE: Dict[tuple, tuple] = {}


def device_g2_bases(ctx, points, n_needed: int):
    """A device G2 set for `points` (a g2.G2Bases or a sequence of G2
    points), made once per context and sequence."""
    from ... import g2 as G2
    hip = _hip(ctx)
    key = (id(hip), id(points))
    hit = _G2_CACHE.get(key)
    n = points.n if isinstance(points, G2.G2Bases) else len(points)
    if hit is not None and hit[0] is points and hit[1].n == n:
        gb = hit[1]
    else:
        words = points.words if isinstance(points, G2.G2Bases) else G2.G2Bases(points).words
        gb = hip.msm_g2_bases_create(words)
        _G2_CACHE[key] = (points, gb)
    if n_needed > gb.n:
        raise ValueError("scalars length exceeds points length")
    return gb


def _msm_g2(ctx):
    """MSM_G2 (reference: runtime/kernels/cpu/kernels.py:338-348 over
    crypto/msm/pippenger.py:67-121): sum s_i Q_i on the device, scalars an
    FR HipBuffer; the result a host G2 point."""
    from ... import g2 as G2
    node = ctx["node"]
    points: Buffer = ctx["inputs"][0]
    scalars: Buffer = ctx["inputs"][1]
    if scalars.dtype != DType.FR:
        raise ValueError("hip msm_g2 scalars must be FR")
    if not isinstance(scalars.data, HipBuffer):
        raise ValueError("hip msm_g2 expects scalars to be a HipBuffer")
    n = int(scalars.data.n)
    gb = device_g2_bases(ctx, points.data, n)
    acc = G2.point_from_words(_hip(ctx).msm_g2(gb, scalars.data.ptr, n, _window_bits(ctx))) if n else G2.G2_ZERO
    oid = node.outputs[0]
    return {"outputs": {oid: Buffer(id=oid, device=Device.CPU, dtype=DType.G2, data=acc)}}
