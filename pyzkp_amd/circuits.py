"""Small R1CS builder and the reference's two example circuits, so that
real circuits run end to end through the HIP Groth16 path.

The reference builds circuits with a gnark-like frontend (`frontend/api`,
`compile_circuit`, `build_witness`, `check_r1cs`); circuit authoring itself
is not part of this package.  `R1CSBuilder` is the minimum needed for the
reference's CubicCircuit (tests/correctness/cpu/test_groth16_e2e.py:11-25)
and Poseidon gadget (frontend/api/std/poseidon.py:14-55): linear
combinations stay symbolic (additions, constants and MDS mixing cost no
constraint), every product of two non-constant combinations is one R1CS
row a * b = c, and an assertion a == b is the row (a - b) * 1 = 0.  The
witness is computed while the circuit is built.  `compile()` returns an
object with the interface the Groth16 setup / prover use (n, n_vars, A, B,
C as CSR, public_ids, aux_ids; the constant one is the last variable), plus
the witness words.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .field import FR_MODULUS, ints_to_words, words_to_ints

ONE = -1  # key of the constant term in a linear combination


class LC(dict):
    """Linear combination {variable id (ONE for the constant): coefficient}."""

    def value(self, vals: List[int]) -> int:
        return sum(c * (1 if v == ONE else vals[v]) for v, c in self.items()) % FR_MODULUS

    def is_const(self) -> bool:
        return all(v == ONE for v in self)


@dataclass
class CompiledCircuit:
    """What setup / prove / the key cache read from a circuit."""
    log_n: int
    A: object
    B: object
    C: object
    m: int                                   # variables before the constant one
    public_ids: List[int] = field(default_factory=list)
    aux_ids: np.ndarray = None
    n_constraints: int = 0

    @property
    def n(self) -> int:
        return 1 << self.log_n

    @property
    def n_vars(self) -> int:
        return self.m + 1

    @property
    def one_id(self) -> int:
        return self.m


class R1CSBuilder:
    def __init__(self) -> None:
        self.vals: List[int] = []
        self.public: List[int] = []
        self.rows: List[Tuple[LC, LC, LC]] = []

    # variables ------------------------------------------------------------
    def _new(self, value: int) -> LC:
        self.vals.append(int(value) % FR_MODULUS)
        return LC({len(self.vals) - 1: 1})

    def secret(self, value: int) -> LC:
        return self._new(value)

    def public_input(self, value: int) -> LC:
        lc = self._new(value)
        self.public.append(len(self.vals) - 1)
        return lc

    @staticmethod
    def constant(c: int) -> LC:
        c = int(c) % FR_MODULUS
        return LC({ONE: c}) if c else LC()

    # operations -----------------------------------------------------------
    @staticmethod
    def _lc(x) -> LC:
        return x if isinstance(x, LC) else R1CSBuilder.constant(x)

    def add(self, *xs) -> LC:
        out = LC()
        for x in xs:
            for v, c in self._lc(x).items():
                s = (out.get(v, 0) + c) % FR_MODULUS
                if s:
                    out[v] = s
                else:
                    out.pop(v, None)
        return out

    def scale(self, x, k: int) -> LC:
        k = int(k) % FR_MODULUS
        return LC({v: c * k % FR_MODULUS for v, c in self._lc(x).items() if c * k % FR_MODULUS})

    def sub(self, a, b) -> LC:
        return self.add(a, self.scale(b, FR_MODULUS - 1))

    def mul(self, *xs) -> LC:
        """Product of the arguments; one constraint per product of two
        non-constant combinations."""
        acc = self._lc(xs[0])
        for x in xs[1:]:
            x = self._lc(x)
            if acc.is_const():
                acc = self.scale(x, acc.get(ONE, 0))
            elif x.is_const():
                acc = self.scale(acc, x.get(ONE, 0))
            else:
                out = self._new(acc.value(self.vals) * x.value(self.vals))
                self.rows.append((acc, x, out))
                acc = out
        return acc

    def assert_equal(self, a, b) -> None:
        self.rows.append((self.sub(a, b), self.constant(1), LC()))

    # output ---------------------------------------------------------------
    def compile(self, min_log_n: int = 1) -> Tuple[CompiledCircuit, np.ndarray]:
        from .groth16.synthetic import CSR
        m = len(self.vals)
        rows = len(self.rows)
        log_n = max(min_log_n, (max(rows, 1) - 1).bit_length())
        n = 1 << log_n

        def csr(k: int) -> CSR:
            rp = np.zeros(n + 1, dtype=np.uint32)
            cols, coefs = [], []
            for j, r in enumerate(self.rows):
                for v, c in sorted(r[k].items()):
                    cols.append(m if v == ONE else v)
                    coefs.append(c)
                rp[j + 1] = len(cols)
            rp[rows + 1:] = len(cols)
            return CSR(rp, np.array(cols, dtype=np.uint32),
                       ints_to_words(coefs) if coefs else np.zeros((0, 4), dtype=np.uint64))

        pub = set(self.public)
        circ = CompiledCircuit(log_n=log_n, A=csr(0), B=csr(1), C=csr(2), m=m, public_ids=[m] + list(self.public),
                               aux_ids=np.array([i for i in range(m) if i not in pub], dtype=np.uint32),
                               n_constraints=rows)
        w = ints_to_words(self.vals + [1])
        return circ, w


def check_r1cs(circ, w_words: np.ndarray) -> None:
    """AssertionError if the witness violates a row (the role of
    frontend/api/witness.py check_r1cs)."""
    w = words_to_ints(w_words)
    for j, (ra, rb, rc) in enumerate(zip(circ.A.row_dicts(), circ.B.row_dicts(), circ.C.row_dicts())):
        ev = lambda r: sum(c * w[v] for v, c in r.items()) % FR_MODULUS
        if ev(ra) * ev(rb) % FR_MODULUS != ev(rc):
            raise AssertionError(f"constraint {j} not satisfied")


# ------------------------------------------------------------------ circuits
def cubic_circuit(x: int, y: Optional[int] = None):
    """y = x^3 + x + 5 with x secret and y public (the reference's
    CubicCircuit); y defaults to the true value."""
    b = R1CSBuilder()
    xv = b.secret(x)
    y = (x * x * x + x + 5) % FR_MODULUS if y is None else y
    yv = b.public_input(y)
    x3 = b.mul(xv, xv, xv)
    b.assert_equal(yv, b.add(x3, xv, 5))
    return b.compile()


def poseidon_constants(t: int, r_f: int, r_p: int) -> Tuple[List[int], List[List[int]]]:
    """Round constants and mixing matrix as the reference derives them
    (crypto/hash/poseidon.py:8-27): SHA-256 of fixed seed strings, reduced
    mod r."""
    h = lambda s: int.from_bytes(hashlib.sha256(s.encode()).digest(), "big") % FR_MODULUS
    arc = [h(f"poseidon_arc_{t}_{r_f}_{r_p}_{i}") for i in range((r_f + r_p) * t)]
    mds = [[h(f"poseidon_mds_{t}_{r_f}_{r_p}_{i}_{j}") for j in range(t)] for i in range(t)]
    return arc, mds


def poseidon_permutation_out(b: R1CSBuilder, inputs: Sequence[LC], r_f: int = 8, r_p: int = 53) -> LC:
    """The Poseidon gadget (frontend/api/std/poseidon.py:14-55): state
    [0, inputs...], per round constants, x^5 on every element in full rounds
    and on element 0 in partial rounds, then the matrix; returns state[0]."""
    t = len(inputs) + 1
    arc, mds = poseidon_constants(t, r_f, r_p)
    state = [b.constant(0)] + list(inputs)
    for r in range(r_f + r_p):
        state = [b.add(state[i], arc[r * t + i]) for i in range(t)]
        full = r < r_f // 2 or r >= r_f // 2 + r_p
        for i in range(t):
            if full or i == 0:
                x2 = b.mul(state[i], state[i])
                x4 = b.mul(x2, x2)
                state[i] = b.mul(x4, state[i])
        state = [b.add(*[b.scale(state[j], mds[i][j]) for j in range(t)]) for i in range(t)]
    return state[0]


def poseidon_hash(inputs: Sequence[int], r_f: int = 8, r_p: int = 53) -> int:
    """The same permutation on plain field elements (crypto/hash/poseidon.py:29-62)."""
    t = len(inputs) + 1
    arc, mds = poseidon_constants(t, r_f, r_p)
    state = [0] + [int(x) % FR_MODULUS for x in inputs]
    for r in range(r_f + r_p):
        state = [(state[i] + arc[r * t + i]) % FR_MODULUS for i in range(t)]
        full = r < r_f // 2 or r >= r_f // 2 + r_p
        state = [pow(s, 5, FR_MODULUS) if (full or i == 0) else s for i, s in enumerate(state)]
        state = [sum(mds[i][j] * state[j] for j in range(t)) % FR_MODULUS for i in range(t)]
    return state[0]


def poseidon_circuit(in0: int, in1: int, expected: Optional[int] = None):
    """Poseidon(in0, in1) == expected, inputs secret and expected public
    (the reference's PoseidonCircuit, tests/correctness/cpu/test_poseidon.py)."""
    b = R1CSBuilder()
    a0, a1 = b.secret(in0), b.secret(in1)
    exp = poseidon_hash([in0, in1]) if expected is None else expected
    ev = b.public_input(exp)
    b.assert_equal(poseidon_permutation_out(b, [a0, a1]), ev)
    return b.compile()
