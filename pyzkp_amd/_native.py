"""ctypes binding of libzkhip.so (the C-ABI in include/zkhip.h).

The library is loaded from this package directory only (it is built in-tree
by ``pyzkp_amd.build``).  There is no fallback: if the library is missing or
no device is present, every entry point raises.
"""
from __future__ import annotations

import ctypes as ct
import os
import threading
from typing import Optional, Sequence

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libzkhip.so")

ZK_OK = 0
ZK_ERR_INVALID = 1
ZK_ERR_RUNTIME = 2
ZK_ERR_NO_DEVICE = 3

_u64p = ct.POINTER(ct.c_uint64)
_vp = ct.c_void_p

# (name, argtypes) for every symbol declared in include/zkhip.h
SIGNATURES = {
    "zk_version": [],
    "zk_device_count": [ct.POINTER(ct.c_int)],
    "zk_init": [ct.c_int, ct.POINTER(_vp)],
    "zk_destroy": [_vp],
    "zk_last_error": [_vp],
    "zk_alloc": [_vp, ct.c_size_t, ct.POINTER(_vp)],
    "zk_free": [_vp, _vp],
    "zk_pinned_alloc": [_vp, ct.c_size_t, ct.POINTER(_vp)],
    "zk_pinned_free": [_vp, _vp],
    "zk_h2d": [_vp, _vp, _vp, ct.c_size_t, _vp],
    "zk_d2h": [_vp, _vp, _vp, ct.c_size_t, _vp],
    "zk_d2d": [_vp, _vp, _vp, ct.c_size_t, _vp],
    "zk_memset_zero": [_vp, _vp, ct.c_size_t, _vp],
    "zk_sync": [_vp, _vp],
    "zk_fr_to_mont": [_vp, _vp, _vp, ct.c_uint64, _vp],
    "zk_fr_from_mont": [_vp, _vp, _vp, ct.c_uint64, _vp],
    "zk_fq_to_mont": [_vp, _vp, _vp, ct.c_uint64, _vp],
    "zk_fq_from_mont": [_vp, _vp, _vp, ct.c_uint64, _vp],
    "zk_fr_mul": [_vp, _vp, _vp, _vp, ct.c_uint64, _vp],
    "zk_fr_add": [_vp, _vp, _vp, _vp, ct.c_uint64, _vp],
    "zk_fr_sub": [_vp, _vp, _vp, _vp, ct.c_uint64, _vp],
    "zk_fr_scale": [_vp, _vp, _u64p, _vp, ct.c_uint64, _vp],
    "zk_ntt": [_vp, _vp, ct.c_uint64, _vp, ct.c_uint32, ct.c_int, _u64p, _u64p, _vp],
    "zk_ntt_batch": [_vp, _vp, _vp, ct.c_uint32, ct.c_uint32, ct.c_int, _u64p, _u64p, _vp],
    "zk_msm_g1": [_vp, _vp, _vp, ct.c_uint64, ct.c_int, _u64p, _vp],
    "zk_msm_g1_noalloc": [_vp, _vp, _vp, ct.c_uint64, ct.c_int, _u64p, _vp],
    "zk_msm_bases_create": [_vp, _vp, ct.c_uint64, ct.POINTER(_vp), _vp],
    "zk_msm_bases_destroy": [_vp, _vp],
    "zk_msm_g1_prepared": [_vp, _vp, _vp, ct.c_uint64, ct.c_int, _u64p, _vp],
    "zk_msm_g2_bases_create": [_vp, _vp, ct.c_uint64, ct.POINTER(_vp), _vp],
    "zk_msm_g2_bases_destroy": [_vp, _vp],
    "zk_msm_g2": [_vp, _vp, _vp, ct.c_uint64, ct.c_int, _u64p, _vp],
    "zk_g1_fixed_base": [_vp, _vp, _vp, ct.c_uint64, _vp, _vp],
    "zk_fixed_bases_create": [_vp, _vp, ct.c_uint64, ct.c_int, ct.POINTER(_vp), _vp],
    "zk_fixed_bases_destroy": [_vp, _vp],
    "zk_fixed_bases_info": [_vp, ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_int)],
    "zk_msm_g1_fixed_batch": [_vp, _vp, _vp, _u64p, _u64p, _u64p, ct.c_uint32, _u64p, _vp],
    "zk_batch_create": [_vp, _vp, _vp, _u64p, _u64p, _u64p, ct.c_uint32, _vp, ct.c_uint32, ct.c_int,
                        ct.POINTER(_vp), _vp],
    "zk_batch_run": [_vp, _vp, _u64p, _vp],
    "zk_batch_destroy": [_vp, _vp],
    "zk_fr_powers": [_vp, _u64p, _u64p, ct.c_uint64, _vp, _vp],
    "zk_synth_witness": [_vp, _vp, ct.c_uint32, ct.c_uint32, ct.c_int, _vp],
    "zk_r1cs_eval": [_vp, _vp, _vp, _vp, _vp, ct.c_uint64, _vp, _vp],
    "zk_fr_gather": [_vp, _vp, _vp, ct.c_uint64, _vp, _vp],
    "zk_fr_inv": [_vp, _vp, _vp, ct.c_uint64, _vp],
    "zk_fr_count_nonzero": [_vp, _vp, ct.c_uint64, ct.POINTER(ct.c_uint64), _vp],
    "zk_plonk_quotient_evals": [_vp, ct.POINTER(ct.c_void_p), ct.c_uint64, ct.c_uint64] + [_vp] * 8
                               + [_vp, _vp, _vp],
    "zk_g2_msm_host": [_vp, _vp, ct.c_uint64, ct.c_int, ct.c_int, _vp],
    "zk_g2_mul_fixed_host": [_vp, _vp, ct.c_uint64, ct.c_int, _vp],
    "zk_g2_add_host": [_vp, _vp, _vp],
    "zk_host_simd": [ct.c_int],
    "zk_g2_bases_create": [_vp, ct.c_uint64, ct.c_int, ct.POINTER(_vp)],
    "zk_g2_bases_destroy": [_vp],
    "zk_g2_msm_prepared": [_vp, _vp, ct.c_uint64, ct.c_int, ct.c_int, _vp],
    "zk_g1_mul_host": [_vp, _vp, _vp],
    "zk_pairing_product": [_vp, _vp, ct.c_uint32, _vp],
    "zk_pairing_check": [_vp, _vp, ct.c_uint32, ct.POINTER(ct.c_int)],
    "zk_set_tuning": [_vp, ct.c_char_p, ct.c_int64],
    "zk_probe_mad_rate": [_vp, ct.POINTER(ct.c_double)],
    "zk_profile_enable": [_vp, ct.c_int],
    "zk_profile_get": [_vp, ct.c_char_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_uint64)],
    "zk_profile_reset": [_vp],
}

class NttJob(ct.Structure):
    """zk_ntt_job of include/zkhip.h."""
    _fields_ = [("in_", ct.c_void_p), ("out", ct.c_void_p), ("log_n", ct.c_uint32), ("batch", ct.c_uint32),
                ("inverse", ct.c_int), ("has_omega", ct.c_int), ("has_shift", ct.c_int),
                ("omega_mont", ct.c_uint64 * 4), ("shift_mont", ct.c_uint64 * 4)]


_lib = None
_lib_lock = threading.Lock()


class ZkError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH) -> ct.CDLL:
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ZkError(f"{path} is not built: run `python -m pyzkp_amd.build` first")
        lib = ct.CDLL(path)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ct.c_char_p if name == "zk_last_error" else ct.c_int
        _lib = lib
        return lib


def native_available() -> bool:
    """True when the library loads and a HIP device is visible."""
    try:
        lib = load_library()
    except (ZkError, OSError):
        return False
    n = ct.c_int(0)
    lib.zk_device_count(ct.byref(n))
    return n.value > 0


def _words(x: int) -> "ct.Array":
    arr = (ct.c_uint64 * 4)()
    for i in range(4):
        arr[i] = (int(x) >> (64 * i)) & 0xFFFFFFFFFFFFFFFF
    return arr


class DeviceBuffer:
    """A device allocation owned by one Context (freed on ``free()``)."""

    __slots__ = ("ctx", "ptr", "nbytes", "__weakref__")

    def __init__(self, ctx: "Context", ptr: int, nbytes: int) -> None:
        self.ctx = ctx
        self.ptr = ptr
        self.nbytes = nbytes

    def free(self) -> None:
        if self.ptr:
            self.ctx._check(self.ctx.lib.zk_free(self.ctx.handle, self.ptr))
            self.ptr = 0

    def offset(self, nbytes: int) -> int:
        return self.ptr + nbytes


class MsmG2Bases:
    """A G2 base set on the device, converted once (zk_msm_g2_bases_*)."""

    def __init__(self, ctx: "Context", handle, n: int) -> None:
        self.ctx = ctx
        self.handle = handle
        self.n = n

    def free(self) -> None:
        if self.handle is not None and self.handle.value:
            self.ctx._check(self.ctx.lib.zk_msm_g2_bases_destroy(self.ctx.handle, self.handle))
            self.handle = None


class MsmBases:
    """A base set converted once for repeated MSMs (zk_msm_bases_*)."""

    def __init__(self, ctx: "Context", handle, n: int) -> None:
        self.ctx = ctx
        self.handle = handle
        self.n = n

    def free(self) -> None:
        if self.handle:
            self.ctx._check(self.ctx.lib.zk_msm_bases_destroy(self.ctx.handle, self.handle))
            self.handle = None


class FixedBases:
    """A base set expanded into fixed-base tables on the device."""

    def __init__(self, ctx: "Context", handle, n: int, window_bits: int) -> None:
        self.ctx = ctx
        self.handle = handle
        self.n = n
        self.window_bits = window_bits

    def free(self) -> None:
        if self.handle:
            self.ctx._check(self.ctx.lib.zk_fixed_bases_destroy(self.ctx.handle, self.handle))
            self.handle = None


class BatchPlan:
    """Fixed-base MSMs + NTT jobs, optionally replayed from one hipGraph."""

    def __init__(self, ctx: "Context", handle, n_msm: int) -> None:
        self.ctx = ctx
        self.handle = handle
        self.n_msm = n_msm

    def run(self, stream=None):
        out = (ct.c_uint64 * (12 * max(self.n_msm, 1)))()
        self.ctx._check(self.ctx.lib.zk_batch_run(self.ctx.handle, self.handle, out, stream))
        w = list(out)
        f = lambda a: w[a] | (w[a + 1] << 64) | (w[a + 2] << 128) | (w[a + 3] << 192)
        return [(f(12 * j), f(12 * j + 4), f(12 * j + 8)) for j in range(self.n_msm)]

    def free(self) -> None:
        if self.handle:
            self.ctx._check(self.ctx.lib.zk_batch_destroy(self.ctx.handle, self.handle))
            self.handle = None


class Context:
    """One HIP device context (zk_init / zk_destroy)."""

    def __init__(self, ordinal: int = 0) -> None:
        self.lib = load_library()
        h = _vp()
        rc = self.lib.zk_init(int(ordinal), ct.byref(h))
        if rc != ZK_OK:
            msg = self.lib.zk_last_error(None).decode()
            raise (ValueError if rc == ZK_ERR_INVALID else RuntimeError)(f"zk_init: {msg}")
        self.handle = h
        self.ordinal = ordinal

    # ---- errors
    def _check(self, rc: int) -> None:
        if rc == ZK_OK:
            return
        msg = self.lib.zk_last_error(self.handle)
        msg = msg.decode() if msg else "unknown error"
        if rc == ZK_ERR_INVALID:
            raise ValueError(msg)
        raise RuntimeError(msg)

    def close(self) -> None:
        if self.handle:
            self.lib.zk_destroy(self.handle)
            self.handle = None

    # ---- memory
    def alloc(self, nbytes: int) -> DeviceBuffer:
        p = _vp()
        self._check(self.lib.zk_alloc(self.handle, int(nbytes), ct.byref(p)))
        return DeviceBuffer(self, p.value, int(nbytes))

    def pinned_alloc(self, nbytes: int) -> int:
        p = _vp()
        self._check(self.lib.zk_pinned_alloc(self.handle, int(nbytes), ct.byref(p)))
        return p.value

    def pinned_free(self, ptr: int) -> None:
        self._check(self.lib.zk_pinned_free(self.handle, ptr))

    def h2d(self, dst: int, src, nbytes: int, stream: Optional[int] = None) -> None:
        self._check(self.lib.zk_h2d(self.handle, dst, src, int(nbytes), stream))

    def d2h(self, dst, src: int, nbytes: int, stream: Optional[int] = None) -> None:
        self._check(self.lib.zk_d2h(self.handle, dst, src, int(nbytes), stream))

    def d2d(self, dst: int, src: int, nbytes: int, stream: Optional[int] = None) -> None:
        self._check(self.lib.zk_d2d(self.handle, dst, src, int(nbytes), stream))

    def memset_zero(self, dst: int, nbytes: int, stream: Optional[int] = None) -> None:
        self._check(self.lib.zk_memset_zero(self.handle, dst, int(nbytes), stream))

    def sync(self, stream: Optional[int] = None) -> None:
        self._check(self.lib.zk_sync(self.handle, stream))

    def upload(self, data: bytes, stream: Optional[int] = None) -> DeviceBuffer:
        buf = self.alloc(max(len(data), 1))
        if data:
            self.h2d(buf.ptr, data, len(data), stream)
            self.sync(stream)
        return buf

    def download(self, src: int, nbytes: int, stream: Optional[int] = None) -> bytes:
        out = ct.create_string_buffer(int(nbytes))
        if nbytes:
            self.d2h(out, src, nbytes, stream)
            self.sync(stream)
        return out.raw

    # ---- element-wise
    def fr_to_mont(self, src: int, dst: int, n: int, stream=None) -> None:
        self._check(self.lib.zk_fr_to_mont(self.handle, src, dst, int(n), stream))

    def fr_from_mont(self, src: int, dst: int, n: int, stream=None) -> None:
        self._check(self.lib.zk_fr_from_mont(self.handle, src, dst, int(n), stream))

    def fq_to_mont(self, src: int, dst: int, n: int, stream=None) -> None:
        self._check(self.lib.zk_fq_to_mont(self.handle, src, dst, int(n), stream))

    def fq_from_mont(self, src: int, dst: int, n: int, stream=None) -> None:
        self._check(self.lib.zk_fq_from_mont(self.handle, src, dst, int(n), stream))

    def fr_mul(self, a: int, b: int, out: int, n: int, stream=None) -> None:
        self._check(self.lib.zk_fr_mul(self.handle, a, b, out, int(n), stream))

    def fr_add(self, a: int, b: int, out: int, n: int, stream=None) -> None:
        self._check(self.lib.zk_fr_add(self.handle, a, b, out, int(n), stream))

    def fr_sub(self, a: int, b: int, out: int, n: int, stream=None) -> None:
        self._check(self.lib.zk_fr_sub(self.handle, a, b, out, int(n), stream))

    def fr_scale(self, a: int, c_mont: int, out: int, n: int, stream=None) -> None:
        self._check(self.lib.zk_fr_scale(self.handle, a, _words(c_mont), out, int(n), stream))

    # ---- NTT / MSM
    def ntt(self, src: int, in_len: int, dst: int, log_n: int, inverse: bool,
            omega_mont: Optional[int] = None, shift_mont: Optional[int] = None, stream=None) -> None:
        w = _words(omega_mont) if omega_mont is not None else None
        s = _words(shift_mont) if shift_mont is not None else None
        self._check(self.lib.zk_ntt(self.handle, src, int(in_len), dst, int(log_n), int(bool(inverse)),
                                    w, s, stream))

    def ntt_batch(self, src: int, dst: int, log_n: int, batch: int, inverse: bool,
                  omega_mont: Optional[int] = None, shift_mont: Optional[int] = None, stream=None) -> None:
        w = _words(omega_mont) if omega_mont is not None else None
        s = _words(shift_mont) if shift_mont is not None else None
        self._check(self.lib.zk_ntt_batch(self.handle, src, dst, int(log_n), int(batch), int(bool(inverse)),
                                          w, s, stream))

    def msm_g1(self, bases: int, scalars: int, n: int, window_bits: int = 0, stream=None) -> Sequence[int]:
        """Returns Jacobian Montgomery (X, Y, Z) as three ints (Z = 0: identity)."""
        out = (ct.c_uint64 * 12)()
        self._check(self.lib.zk_msm_g1(self.handle, bases, scalars, int(n), int(window_bits), out, stream))
        w = list(out)
        f = lambda a: w[a] | (w[a + 1] << 64) | (w[a + 2] << 128) | (w[a + 3] << 192)
        return f(0), f(4), f(8)

    def msm_bases_create(self, bases: int, n: int, stream=None) -> "MsmBases":
        h = _vp()
        self._check(self.lib.zk_msm_bases_create(self.handle, bases, int(n), ct.byref(h), stream))
        return MsmBases(self, h, int(n))

    def msm_g2_bases_create(self, words, stream=None) -> "MsmG2Bases":
        """Upload a G2 base set: (n, 16) uint64 host words, affine canonical
        plain (x0, x1, y0, y1), zeros for the identity."""
        import numpy as np
        w = np.ascontiguousarray(words, dtype="<u8")
        if w.ndim != 2 or w.shape[1] != 16:
            raise ValueError(f"G2 base words must have shape (n, 16), got {tuple(w.shape)}")
        h = _vp()
        self._check(self.lib.zk_msm_g2_bases_create(self.handle, w.ctypes.data, int(w.shape[0]), ct.byref(h), stream))
        return MsmG2Bases(self, h, int(w.shape[0]))

    def msm_g2(self, gb: "MsmG2Bases", scalars: int, n: int, window_bits: int = 0, stream=None):
        """sum_i s_i Q_i over the first n bases of a device G2 set, scalars on
        the device (Fr Montgomery); the affine canonical result as 16 uint64
        words (zeros for the identity), the layout zk_g2_msm_host writes."""
        import numpy as np
        out = np.zeros(16, dtype="<u8")
        self._check(self.lib.zk_msm_g2(self.handle, gb.handle, scalars, int(n), int(window_bits),
                                       out.ctypes.data_as(_u64p), stream))
        return out

    def msm_g1_prepared(self, mb: "MsmBases", scalars: int, n: int, window_bits: int = 0,
                        stream=None) -> Sequence[int]:
        """zk_msm_g1 over the first n bases of a prepared set."""
        out = (ct.c_uint64 * 12)()
        self._check(self.lib.zk_msm_g1_prepared(self.handle, mb.handle, scalars, int(n), int(window_bits), out,
                                                stream))
        w = list(out)
        f = lambda a: w[a] | (w[a + 1] << 64) | (w[a + 2] << 128) | (w[a + 3] << 192)
        return f(0), f(4), f(8)

    def fr_powers(self, base_mont: int, scale_mont: int, n: int, out: int, stream=None) -> None:
        self._check(self.lib.zk_fr_powers(self.handle, _words(base_mont), _words(scale_mont), int(n), out, stream))

    def r1cs_eval(self, row_ptr: int, cols: int, coeffs: int, w: int, rows: int, out: int, stream=None) -> None:
        self._check(self.lib.zk_r1cs_eval(self.handle, row_ptr, cols, coeffs, w, int(rows), out, stream))

    def fr_gather(self, src: int, idx: int, n: int, out: int, stream=None) -> None:
        self._check(self.lib.zk_fr_gather(self.handle, src, idx, int(n), out, stream))

    def fr_inv(self, a: int, out: int, n: int, stream=None) -> None:
        self._check(self.lib.zk_fr_inv(self.handle, a, out, int(n), stream))

    def plonk_quotient_evals(self, ext, n: int, m: int, omega_m_mont: int, shift_mont: int, shift_n_mont: int,
                             alpha_mont: int, beta_mont: int, gamma_mont: int, k1_mont: int, k2_mont: int,
                             num_out: int, zh_out: int, stream=None) -> None:
        """Fused PLONK quotient numerator / Z_H over the size-m coset; ext = 15
        device pointers (a, b, c, z, pi, ql, qr, qm, qo, qc, s1, s2, s3,
        z(omega X), L1), scalars as Montgomery ints."""
        if len(ext) != 15:
            raise ValueError("plonk quotient needs 15 evaluation vectors")
        arr = (ct.c_void_p * 15)(*[int(p) for p in ext])
        w = [_words(v) for v in (omega_m_mont, shift_mont, shift_n_mont, alpha_mont, beta_mont, gamma_mont,
                                 k1_mont, k2_mont)]
        self._check(self.lib.zk_plonk_quotient_evals(self.handle, arr, int(n), int(m), *w, num_out, zh_out,
                                                     stream))

    def fr_count_nonzero(self, a: int, n: int, stream=None) -> int:
        r = ct.c_uint64(0)
        self._check(self.lib.zk_fr_count_nonzero(self.handle, a, int(n), ct.byref(r), stream))
        return r.value

    def g1_fixed_base(self, table: int, scalars: int, n: int, out: int, stream=None) -> None:
        self._check(self.lib.zk_g1_fixed_base(self.handle, table, scalars, int(n), out, stream))

    def fixed_bases_create(self, bases: int, n: int, window_bits: int = 0, stream=None) -> "FixedBases":
        h = _vp()
        self._check(self.lib.zk_fixed_bases_create(self.handle, bases, int(n), int(window_bits), ct.byref(h),
                                                   stream))
        nn = ct.c_uint64(0)
        cb = ct.c_int(0)
        self.lib.zk_fixed_bases_info(h, ct.byref(nn), ct.byref(cb))
        return FixedBases(self, h, nn.value, cb.value)

    def msm_g1_fixed_batch(self, fb: "FixedBases", scalars: int, offsets, sizes, base_offsets=None,
                           stream=None):
        """Returns a list of Jacobian Montgomery (X, Y, Z) int triples."""
        count = len(sizes)
        off = (ct.c_uint64 * max(count, 1))(*[int(x) for x in offsets])
        sz = (ct.c_uint64 * max(count, 1))(*[int(x) for x in sizes])
        bo = (ct.c_uint64 * max(count, 1))(*[int(x) for x in base_offsets]) if base_offsets is not None else None
        out = (ct.c_uint64 * (12 * max(count, 1)))()
        self._check(self.lib.zk_msm_g1_fixed_batch(self.handle, fb.handle, scalars, off, sz, bo, count, out,
                                                   stream))
        w = list(out)
        f = lambda a: w[a] | (w[a + 1] << 64) | (w[a + 2] << 128) | (w[a + 3] << 192)
        return [(f(12 * j), f(12 * j + 4), f(12 * j + 8)) for j in range(count)]

    def batch_create(self, fb: Optional["FixedBases"], scalars: int, offsets, sizes, ntt_jobs,
                     use_graph: bool = True, base_offsets=None, stream=None) -> "BatchPlan":
        """ntt_jobs: list of dicts {in, out, log_n, batch, inverse, omega_mont?, shift_mont?}."""
        n_msm = len(sizes)
        off = (ct.c_uint64 * max(n_msm, 1))(*[int(x) for x in offsets])
        sz = (ct.c_uint64 * max(n_msm, 1))(*[int(x) for x in sizes])
        jobs = (NttJob * max(len(ntt_jobs), 1))()
        for k, j in enumerate(ntt_jobs):
            jobs[k].in_ = j["in"]
            jobs[k].out = j["out"]
            jobs[k].log_n = j["log_n"]
            jobs[k].batch = j.get("batch", 1)
            jobs[k].inverse = int(bool(j.get("inverse", False)))
            if j.get("omega_mont") is not None:
                jobs[k].has_omega = 1
                jobs[k].omega_mont = _words(j["omega_mont"])
            if j.get("shift_mont") is not None:
                jobs[k].has_shift = 1
                jobs[k].shift_mont = _words(j["shift_mont"])
        bo = (ct.c_uint64 * max(n_msm, 1))(*[int(x) for x in base_offsets]) if base_offsets is not None else None
        h = _vp()
        self._check(self.lib.zk_batch_create(self.handle, fb.handle if fb else None, scalars, off, sz, bo, n_msm,
                                             jobs, len(ntt_jobs), int(bool(use_graph)), ct.byref(h), stream))
        return BatchPlan(self, h, n_msm)

    # ---- tuning / profiling
    def set_tuning(self, key: str, value: int) -> None:
        self._check(self.lib.zk_set_tuning(self.handle, key.encode(), int(value)))

    def probe_mad_rate(self) -> float:
        r = ct.c_double(0.0)
        self._check(self.lib.zk_probe_mad_rate(self.handle, ct.byref(r)))
        return r.value

    def profile_enable(self, on: bool = True) -> None:
        self._check(self.lib.zk_profile_enable(self.handle, int(bool(on))))

    def profile_get(self, phase: str):
        ms = ct.c_double(0.0)
        cnt = ct.c_uint64(0)
        self._check(self.lib.zk_profile_get(self.handle, phase.encode(), ct.byref(ms), ct.byref(cnt)))
        return ms.value, cnt.value

    def profile_reset(self) -> None:
        self._check(self.lib.zk_profile_reset(self.handle))


_default_ctx: dict = {}


def get_context(ordinal: int = 0) -> Context:
    """Process-wide context per device (created on first use)."""
    ctx = _default_ctx.get(ordinal)
    if ctx is None:
        ctx = Context(ordinal)
        _default_ctx[ordinal] = ctx
    return ctx
