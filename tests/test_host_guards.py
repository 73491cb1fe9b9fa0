"""Host-side guards, no GPU: the prover workspace follows its key, short
witnesses are refused before any host or device read, a key file whose
arrays do not fit its circuit is refused before anything is uploaded, and
the pool frees its pinned staging area."""
import json

import numpy as np
import pytest

from pyzkp_amd import g2 as G2
from pyzkp_amd.groth16 import synthetic_circuit
import importlib

P = importlib.import_module("pyzkp_amd.groth16.prove")
from pyzkp_amd.groth16.keycache import CacheMismatchError, circuit_fingerprint, load_key
from pyzkp_amd.groth16.types import ProvingKey
from pyzkp_amd.runtime import MemoryPool


class _Buf:
    def __init__(self, log, nbytes):
        self.ptr, self.nbytes, self.log = 0x1000 + len(log), nbytes, log
        log.append(("alloc", nbytes))

    def free(self):
        self.log.append(("free", self.nbytes))
        self.ptr = 0


class _Ctx:
    """Records allocations; no device."""
    def __init__(self):
        self.log = []

    def alloc(self, nbytes):
        return _Buf(self.log, nbytes)

    def sync(self, stream=None):
        self.log.append(("sync",))

    def pinned_alloc(self, nbytes):
        self.log.append(("pinned_alloc", nbytes))
        return 0xABC000

    def pinned_free(self, ptr):
        self.log.append(("pinned_free", ptr))


def _key(ctx, n, n_vars, n_aux):
    return ProvingKey(vk=None, beta_g1=None, delta_g1=None, aux_ids=np.arange(n_aux, dtype=np.uint32),
                      trapdoor_tau=0, trapdoor_alpha=0, trapdoor_beta=0, trapdoor_gamma=0, trapdoor_delta=0,
                      n=n, n_vars=n_vars, ctx=ctx)


def test_workspace_lives_on_its_key_and_is_freed_with_it():
    ctx = _Ctx()
    small = _key(ctx, 8, 5, 2)
    w1 = P._work(small, 2)
    assert small.work is w1 and w1.k == 2 and w1.n == 8
    assert P._work(small, 1) is w1                   # reused while it fits
    small.free()
    assert small.work is None and w1.S.ptr == 0       # freed with the key
    big = _key(ctx, 64, 40, 10)
    w2 = P._work(big, 1)
    assert w2 is not w1 and w2.n == 64 and w2.seg == 40 + 64 + 10
    assert ("alloc", 32 * w2.seg) in ctx.log
    # a workspace sized for another shape is rebuilt, not reused
    big.n_vars = 41
    w3 = P._work(big, 1)
    assert w3 is not w2 and w3.seg == 41 + 64 + 10 and w2.S.ptr == 0


def test_short_witness_refused():
    ctx = _Ctx()
    pk = _key(ctx, 8, 5, 2)
    with pytest.raises(ValueError, match="witness must have shape"):
        P.check_witness(pk, np.zeros((4, 4), dtype=np.uint64))
    with pytest.raises(ValueError, match="witness must have shape"):
        P.prove_many(pk, [np.zeros((5, 3), dtype=np.uint64)], [(1, 2)])
    bases = G2.from_words(np.zeros((6, 16), dtype=np.uint64))
    with pytest.raises(ValueError, match="scalar words"):
        G2.g2_msm_words(bases, np.zeros((3, 4), dtype=np.uint64), 6)
    assert ctx.log == []                              # nothing allocated or launched


def _meta(circ):
    n, nv, na = circ.n, circ.n_vars, len(circ.aux_ids)
    return {"format": "setup_cache", "version": 1, "scheme": "groth16",
            "field_modulus": "21888242871839275222246405745257275088548364400416034343698204186575808495617",
            "ir_fingerprint": circuit_fingerprint(circ), "n": n, "n_vars": nv,
            "query_offsets": [0, nv, 2 * nv, 2 * nv + n - 1], "query_total": 2 * nv + n - 1 + na,
            "public_ids": [int(i) for i in circ.public_ids], "public_names": ["x"] * len(circ.public_ids)}


def _arrays(circ, meta):
    nv, npub = circ.n_vars, len(circ.public_ids)
    return {"query_points": np.zeros(8 * meta["query_total"], dtype="<u8"),
            "b_query_g2": np.zeros((nv, 16), dtype="<u8"), "aux_ids": np.asarray(circ.aux_ids, dtype="<u4"),
            "g1": np.zeros((3, 2, 4), dtype="<u8"), "g2": np.zeros((3, 16), dtype="<u8"),
            "ic": np.zeros((npub, 2, 4), dtype="<u8"), "trapdoor": np.zeros((5, 4), dtype="<u8")}


@pytest.mark.parametrize("edit", ["truncated_points", "query_total", "g2_rows", "aux_ids", "offsets"])
def test_key_file_arrays_checked(tmp_path, edit):
    circ = synthetic_circuit(6, 0x10)
    meta = _meta(circ)
    arr = _arrays(circ, meta)
    if edit == "truncated_points":
        arr["query_points"] = arr["query_points"][:-8]
    elif edit == "query_total":
        meta["query_total"] += 1
    elif edit == "g2_rows":
        arr["b_query_g2"] = arr["b_query_g2"][:-1]
    elif edit == "aux_ids":
        arr["aux_ids"] = arr["aux_ids"].copy()
        arr["aux_ids"][0] += 1
    elif edit == "offsets":
        meta["query_offsets"][3] += 1
    path = tmp_path / "k.npz"
    np.savez(path, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **arr)
    with pytest.raises(CacheMismatchError, match="arrays mismatch"):
        load_key(str(path), circ, ctx=None)          # refused before any upload


def test_pool_close_frees_staging():
    ctx = _Ctx()
    pool = MemoryPool()
    ptr = pool.staging(ctx, 100)
    assert pool.staging(ctx, 50) == ptr
    pool.close()
    assert ("pinned_free", ptr) in ctx.log
    assert pool._staging == {}
