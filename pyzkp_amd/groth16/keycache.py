"""On-disk cache of Groth16 proving keys (the reference's setup cache,
runtime/cache.py:13-97, in a format that executes nothing on load).

A key file is one `.npz` archive written atomically (temporary file, then
rename, as `_save_pickle` does): plain integer arrays for the G1 query
points (affine Montgomery words, as on the device), b_query_g2 (canonical
G2 words), the A / B / C matrices (CSR), the aux ids and the verifying-key
points, plus a JSON meta record.  It is read back with
`numpy.load(allow_pickle=False)`.  The meta checks follow
`load_setup_cache` (format, version, scheme, field modulus, circuit
fingerprint) and raise `CacheMismatchError`, a `ValueError`, with the same
messages.  Loading re-uploads the queries and rebuilds the fixed-base
tables and matrices on the device, so a server restart skips the setup's
host G2 work.
"""
from __future__ import annotations

import hashlib
import json
import os
from typing import Optional

import numpy as np

from ..field import FR_MODULUS, ints_to_words, words_to_ints
from .. import g2 as G2
from .synthetic import SyntheticCircuit
from .types import ProvingKey, VerifyingKey

KEY_CACHE_FORMAT_VERSION = 1
_FORMAT = "setup_cache"
_SCHEME = "groth16"


class CacheMismatchError(ValueError):
    pass


def circuit_fingerprint(circ: SyntheticCircuit) -> str:
    """sha256 over the field modulus, the shape, the public / aux ids and the
    three matrices (the role of circuit_ir_fingerprint, runtime/cache.py:38-50);
    16 hex digits like the reference's."""
    h = hashlib.sha256()
    h.update(json.dumps({"field_modulus": FR_MODULUS, "n": circ.n, "n_vars": circ.n_vars,
                         "public": [int(i) for i in circ.public_ids]}, sort_keys=True).encode())
    h.update(np.ascontiguousarray(circ.aux_ids, dtype="<u4").tobytes())
    for m in (circ.A, circ.B, circ.C):
        for arr, dt in ((m.row_ptr, "<u4"), (m.cols, "<u4"), (m.coeffs, "<u8")):
            h.update(np.ascontiguousarray(arr, dtype=dt).tobytes())
    return h.hexdigest()[:16]


def default_key_path(cache_dir: str, circ: SyntheticCircuit) -> str:
    return os.path.join(cache_dir, "setup", _SCHEME,
                        f"ir_{circuit_fingerprint(circ)}_mod{FR_MODULUS}_v{KEY_CACHE_FORMAT_VERSION}.npz")


def _g1_words(p) -> np.ndarray:
    from ..curve import normalize
    a = normalize(p)
    return np.zeros((2, 4), dtype="<u8") if a is None else ints_to_words([a[0], a[1]])


def _g1_from(w: np.ndarray):
    from ..curve import from_affine
    x, y = words_to_ints(w)
    return from_affine(None if x == 0 and y == 0 else (x, y))


def _g2_words(q) -> np.ndarray:
    return G2._pack([q]).reshape(16)


def save_key(path: str, pk: ProvingKey, circ: SyntheticCircuit) -> dict:
    """Writes pk (with its circuit's fingerprint) to `path`; returns the meta."""
    meta = {"format": _FORMAT, "version": KEY_CACHE_FORMAT_VERSION, "scheme": _SCHEME,
            "field_modulus": str(FR_MODULUS), "ir_fingerprint": circuit_fingerprint(circ),
            "n": pk.n, "n_vars": pk.n_vars, "query_offsets": list(pk.query_offsets),
            "query_total": pk.query_total, "public_ids": [int(i) for i in pk.public_ids],
            "public_names": list(pk.vk.public_names)}
    pts = np.frombuffer(pk.ctx.download(pk.query_points.ptr, 64 * pk.query_total), dtype="<u8")
    vk = pk.vk
    arrays = {
        "meta": np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8),
        "query_points": pts,
        "b_query_g2": np.ascontiguousarray(pk.b_query_g2.words, dtype="<u8"),
        "aux_ids": np.ascontiguousarray(pk.aux_ids, dtype="<u4"),
        "g1": np.stack([_g1_words(p) for p in (vk.alpha_g1, pk.beta_g1, pk.delta_g1)]),
        "g2": np.stack([_g2_words(q) for q in (vk.beta_g2, vk.gamma_g2, vk.delta_g2)]),
        "ic": np.stack([_g1_words(p) for p in vk.ic]),
        "trapdoor": ints_to_words([pk.trapdoor_tau, pk.trapdoor_alpha, pk.trapdoor_beta, pk.trapdoor_gamma,
                                   pk.trapdoor_delta]),
    }
    for name, (rp, cols, co) in pk.csr_host.items():
        arrays[f"{name}_row_ptr"] = np.ascontiguousarray(rp, dtype="<u4")
        arrays[f"{name}_cols"] = np.ascontiguousarray(cols, dtype="<u4")
        arrays[f"{name}_coeffs"] = np.ascontiguousarray(co, dtype="<u8")
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = f"{path}.tmp.npz"
    np.savez(tmp, **arrays)
    os.replace(tmp, path)
    return meta


def _check_arrays(meta: dict, arr: dict, circ: SyntheticCircuit) -> int:
    """Every array the device or the host MSMs will read must have the shape
    the circuit implies; returns the number of G1 query points."""
    def need(cond: bool, what: str) -> None:
        if not cond:
            raise CacheMismatchError(f"setup cache arrays mismatch: {what}")

    n, nv, na = circ.n, circ.n_vars, len(circ.aux_ids)
    total = 2 * nv + (n - 1) + na
    need(int(meta.get("n", -1)) == n and int(meta.get("n_vars", -1)) == nv, "domain or variable count")
    need(int(meta.get("query_total", -1)) == total, "query_total")
    need(list(meta.get("query_offsets", [])) == [0, nv, 2 * nv, 2 * nv + n - 1], "query_offsets")
    need([int(i) for i in meta.get("public_ids", [])] == [int(i) for i in circ.public_ids], "public_ids")
    for name, shape in (("query_points", (8 * total,)), ("b_query_g2", (nv, 16)), ("g1", (3, 2, 4)),
                        ("g2", (3, 16)), ("ic", (len(circ.public_ids), 2, 4)), ("trapdoor", (5, 4))):
        need(name in arr and tuple(arr[name].shape) == shape, name)
    need("aux_ids" in arr and np.array_equal(arr["aux_ids"].astype(np.int64),
                                            np.asarray(circ.aux_ids, dtype=np.int64)), "aux_ids")
    need(len(meta.get("public_names", [])) == len(circ.public_ids), "public_names")
    return total


def load_key(path: str, circ: SyntheticCircuit, ctx, window_bits: int = 0) -> ProvingKey:
    """Reads a key written by save_key for this circuit and rebuilds its
    device state; CacheMismatchError when the file is for something else."""
    from .setup import _upload_csr, key_hbm_bytes
    from .synthetic import CSR
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode())
        if meta.get("format") != _FORMAT:
            raise CacheMismatchError("setup cache format mismatch")
        if int(meta.get("version", -1)) != KEY_CACHE_FORMAT_VERSION:
            raise CacheMismatchError("setup cache version mismatch")
        if str(meta.get("scheme")) != _SCHEME:
            raise CacheMismatchError("setup cache scheme mismatch")
        if int(meta.get("field_modulus", -1)) != FR_MODULUS:
            raise CacheMismatchError("setup cache modulus mismatch")
        if str(meta.get("ir_fingerprint")) != circuit_fingerprint(circ):
            raise CacheMismatchError("setup cache circuit mismatch")
        arr = {k: z[k] for k in z.files}
    total = _check_arrays(meta, arr, circ)
    pts = ctx.upload(arr["query_points"].tobytes() or b"\0" * 64)
    fixed = ctx.fixed_bases_create(pts.ptr, total, window_bits)
    # the matrices come from the circuit the fingerprint was checked against,
    # never from the file
    csr_host = {name: (m.row_ptr, m.cols, m.coeffs) for name, m in (("A", circ.A), ("B", circ.B), ("C", circ.C))}
    csr = {name: _upload_csr(ctx, CSR(*v)) for name, v in csr_host.items()}
    aux = np.ascontiguousarray(circ.aux_ids, dtype="<u4")
    g1 = [_g1_from(w) for w in arr["g1"]]
    g2 = [G2.point_from_words(w) for w in arr["g2"]]
    vk = VerifyingKey(alpha_g1=g1[0], beta_g2=g2[0], gamma_g2=g2[1], delta_g2=g2[2],
                      ic=[_g1_from(w) for w in arr["ic"]], public_names=list(meta["public_names"]))
    td = words_to_ints(arr["trapdoor"])
    pk = ProvingKey(vk=vk, beta_g1=g1[1], delta_g1=g1[2], aux_ids=aux, trapdoor_tau=td[0], trapdoor_alpha=td[1],
                    trapdoor_beta=td[2], trapdoor_gamma=td[3], trapdoor_delta=td[4], n=int(meta["n"]),
                    n_vars=int(meta["n_vars"]), public_ids=list(meta["public_ids"]), ctx=ctx, fixed=fixed,
                    query_offsets=tuple(meta["query_offsets"]), query_points=pts,
                    b_query_g2=G2.from_words(arr["b_query_g2"]), csr=csr,
                    aux_dev=ctx.upload(np.ascontiguousarray(aux, dtype="<u4").tobytes() or b"\0" * 4),
                    query_total=total, csr_host=csr_host)
    pk.hbm_bytes = key_hbm_bytes(pk)
    return pk
