"""Groth16 keys and proofs (reference: protocols/groth16/types.py:9-41),
with the proving key's query vectors resident on the device."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, List, Optional


@dataclass(frozen=True)
class Proof:
    a: Any  # G1 (homogeneous (x, y, 1) ints)
    b: Any  # G2 ((c0, c1) x3)
    c: Any  # G1

    def to_bytes(self) -> bytes:
        """A || B || C as affine big-endian coordinates, 32 bytes each: G1 as
        x || y (the transcript encoding, protocols/plonk/transcript.py:13-19),
        G2 as x.c0 || x.c1 || y.c0 || y.c1; the identity is all zeros."""
        from ..curve import normalize
        from ..g2 import normalize as normalize2

        def be(v: int) -> bytes:
            return int(v).to_bytes(32, "big")

        a, c, b = normalize(self.a), normalize(self.c), normalize2(self.b)
        g1 = lambda p: bytes(64) if p is None else be(p[0]) + be(p[1])
        g2 = bytes(128) if b is None else be(b[0][0]) + be(b[0][1]) + be(b[1][0]) + be(b[1][1])
        return g1(a) + g2 + g1(c)


@dataclass(frozen=True)
class VerifyingKey:
    alpha_g1: Any
    beta_g2: Any
    gamma_g2: Any
    delta_g2: Any
    ic: List[Any]
    public_names: List[str]


@dataclass
class ProvingKey:
    """Reference fields (protocols/groth16/types.py:27-41) plus the device
    state.  The G1 query vectors live in one device fixed-base set laid out
    [a_query | b_query_g1 | h_query | l_query]; `query_offsets` gives each
    start.  b_query_g2 is a packed host copy for the CPU G2 MSM;
    b_query_g2_dev its device base set for the device G2 MSM, made at the
    first device G2 proof."""
    vk: VerifyingKey
    beta_g1: Any
    delta_g1: Any
    aux_ids: Any
    trapdoor_tau: int
    trapdoor_alpha: int
    trapdoor_beta: int
    trapdoor_gamma: int
    trapdoor_delta: int
    n: int = 0
    n_vars: int = 0
    public_ids: List[int] = field(default_factory=list)
    # device state
    ctx: Any = None
    fixed: Any = None                     # FixedBases over all G1 queries
    query_offsets: tuple = ()             # (a, b_g1, h, l) starts in `fixed`
    query_points: Any = None              # DeviceBuffer of the G1 queries (affine R=2^256)
    b_query_g2: Any = None                # g2.G2Bases
    b_query_g2_dev: Any = None            # _native.MsmG2Bases (device), or None
    csr: Any = None                       # dict name -> (row_ptr, cols, coeffs) device buffers
    aux_dev: Any = None                   # DeviceBuffer of aux ids (uint32)
    hbm_bytes: int = 0

    def free(self) -> None:
        if self.work is not None:         # the prover's per-key workspace
            self.work.ctx.sync()
            self.work.free()
            self.work = None
        if self.fixed is not None:
            self.fixed.free()
            self.fixed = None
        for b in [self.query_points, self.aux_dev] + [x for t in (self.csr or {}).values() for x in t]:
            if b is not None:
                b.free()
        self.query_points = self.aux_dev = None
        self.csr = None
        if self.b_query_g2_dev is not None:
            self.b_query_g2_dev.free()
            self.b_query_g2_dev = None
        if self.b_query_g2 is not None and hasattr(self.b_query_g2, "free"):
            self.b_query_g2.free()        # its prepared host set; the words stay

    def g1_query(self, which: str) -> List[Any]:
        """Download one query vector as host points (for small keys / tests)."""
        from ..field import fq_from_mont, unpack_ints
        names = ("a", "b_g1", "h", "l")
        k = names.index(which)
        start = self.query_offsets[k]
        end = self.query_offsets[k + 1] if k + 1 < 4 else self.query_total
        raw = unpack_ints(self.ctx.download(self.query_points.ptr + 64 * start, 64 * (end - start)))
        out = []
        for i in range(0, len(raw), 2):
            x, y = raw[i], raw[i + 1]
            out.append((1, 1, 0) if x == 0 and y == 0 else (fq_from_mont(x), fq_from_mont(y), 1))
        return out

    query_total: int = 0
    csr_host: Any = None                  # dict name -> (row_ptr, cols, coeffs) host arrays
    work: Any = field(default=None, repr=False)   # prover workspace sized for this key
