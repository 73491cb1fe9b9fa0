"""Synthetic R1CS for the served-prover benchmark (no counterpart in the
reference, whose circuits come from its frontend).

`chains` independent chains of `rounds` rounds over a 3-lane state; a round
is the S-box x -> x^5 on each lane (x2 = x x, x4 = x2 x2, x5 = x4 x) and a
3-term linear mix y_k = sum_j M[k][j] x5_j, written as (sum) * 1 = y_k, so
every row has 1-3 non-zeros.  12 constraints per round; the constraint
count is padded with empty rows to n = 2^log_n as compile_r1cs does
(protocols/groth16/r1cs.py:43, 59-62).  Variable ids follow csrc/synth.hip;
the constant one is the last id (one_id = m).  Public: the constant and
lane 0 of chain 0's final state.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from .. import workload
from .._native import load_library
from ..field import words_to_ints

ONE_WORDS = np.array([1, 0, 0, 0], dtype=np.uint64)


@dataclass
class CSR:
    row_ptr: np.ndarray   # uint32 (rows + 1)
    cols: np.ndarray      # uint32 (nnz)
    coeffs: np.ndarray    # uint64 (nnz, 4), canonical

    def transpose(self, ncols: int) -> "CSR":
        rows = np.repeat(np.arange(len(self.row_ptr) - 1, dtype=np.uint32), np.diff(self.row_ptr))
        order = np.argsort(self.cols, kind="stable")
        cnt = np.bincount(self.cols, minlength=ncols).astype(np.uint32)
        rp = np.zeros(ncols + 1, dtype=np.uint32)
        np.cumsum(cnt, out=rp[1:])
        return CSR(rp, rows[order].astype(np.uint32), self.coeffs[order])

    def row_dicts(self) -> List[Dict[int, int]]:
        vals = words_to_ints(self.coeffs) if len(self.cols) else []
        out = []
        for j in range(len(self.row_ptr) - 1):
            a, b = int(self.row_ptr[j]), int(self.row_ptr[j + 1])
            out.append({int(self.cols[k]): vals[k] for k in range(a, b)})
        return out


@dataclass
class SyntheticCircuit:
    log_n: int
    chains: int
    rounds: int
    mix: np.ndarray                 # (9, 4) canonical
    A: CSR
    B: CSR
    C: CSR
    m: int                          # IR variables (one_id = m)
    public_ids: List[int] = field(default_factory=list)
    aux_ids: np.ndarray = None

    @property
    def n(self) -> int:
        return 1 << self.log_n

    @property
    def n_vars(self) -> int:
        return self.m + 1

    @property
    def one_id(self) -> int:
        return self.m


def synthetic_circuit(log_n: int, seed: int) -> SyntheticCircuit:
    n = 1 << log_n
    total_rounds = max(1, n // 12)
    chains = min(8, total_rounds)
    rounds = total_rounds // chains
    m = 3 * chains + 12 * chains * rounds
    one = m
    mix = workload.fr_uniform(seed, 9)
    c = np.arange(chains, dtype=np.int64)[:, None]
    r = np.arange(rounds, dtype=np.int64)[None, :]
    base = 3 * chains + 12 * (c * rounds + r)                       # (chains, rounds)
    prev = np.where(r == 0, 3 * c, base - 12 + 9)                   # lane-0 id of the state entering the round
    nr = chains * rounds * 12
    # per round rows: k = 0..2 : x2_k, x4_k, x5_k (lane-major groups), then y_k
    A_cols, A_cnt, B_cols, C_cols = [], [], [], []
    A_coef = []
    for k in range(3):  # x2_k = x_k x_k
        A_cols.append((prev + k)[..., None])
        B_cols.append(prev + k)
        C_cols.append(base + k)
    for k in range(3):  # x4_k = x2_k x2_k
        A_cols.append((base + k)[..., None])
        B_cols.append(base + k)
        C_cols.append(base + 3 + k)
    for k in range(3):  # x5_k = x4_k x_k
        A_cols.append((base + 3 + k)[..., None])
        B_cols.append(prev + k)
        C_cols.append(base + 6 + k)
    # assemble in row order: rows 0..8 one-term A rows, rows 9..11 mixes
    shape = (chains, rounds)
    one_term = np.stack([a[..., 0] for a in A_cols], axis=-1)        # (chains, rounds, 9)
    mix_cols = np.stack([base + 6, base + 7, base + 8], axis=-1)      # x5_0..2
    # A: 9 rows with 1 nnz + 3 rows with 3 nnz -> 18 nnz per round
    a_cols = np.concatenate([one_term, np.repeat(mix_cols[..., None, :], 3, axis=-2).reshape(*shape, 9)],
                            axis=-1).reshape(-1)
    a_rp_round = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 18], dtype=np.int64)
    a_rp = (np.arange(chains * rounds, dtype=np.int64)[:, None] * 18 + a_rp_round[None, :-1]).reshape(-1)
    a_coef = np.zeros((chains * rounds, 18, 4), dtype=np.uint64)
    a_coef[:, :9] = ONE_WORDS
    a_coef[:, 9:] = mix[None, :, :]
    b_cols = np.concatenate([np.stack(B_cols, axis=-1), np.full(shape + (3,), one)], axis=-1).reshape(-1)
    c_cols = np.concatenate([np.stack(C_cols, axis=-1), base[..., None] + 9 + np.arange(3)], axis=-1).reshape(-1)

    def pad(rp_body, cols, coefs):
        rp = np.empty(n + 1, dtype=np.uint32)
        rp[:nr] = rp_body
        rp[nr:] = len(cols)
        return CSR(rp, cols.astype(np.uint32), coefs)

    A = pad(a_rp, a_cols, a_coef.reshape(-1, 4))
    ones = np.tile(ONE_WORDS, (nr, 1))
    B = pad(np.arange(nr, dtype=np.uint32), b_cols, ones.copy())
    Cm = pad(np.arange(nr, dtype=np.uint32), c_cols, ones.copy())
    out_id = 3 * chains + 12 * (rounds - 1) + 9                      # chain 0, last round, y_0
    pub = [out_id]
    aux = np.array([i for i in range(m) if i not in set(pub)], dtype=np.uint32)
    return SyntheticCircuit(log_n, chains, rounds, mix, A, B, Cm, m, public_ids=[one] + pub, aux_ids=aux)


def synthetic_witness(circ: SyntheticCircuit, seed: int, threads: int = 0) -> np.ndarray:
    """(n_vars, 4) canonical words; the last entry is the constant 1."""
    init = workload.fr_uniform(seed, 3 * circ.chains)
    w = np.zeros((circ.n_vars, 4), dtype=np.uint64)
    mix = np.ascontiguousarray(circ.mix, dtype="<u8")
    rc = load_library().zk_synth_witness(mix.ctypes.data, init.ctypes.data, circ.chains, circ.rounds, threads,
                                         w.ctypes.data)
    if rc:
        raise ValueError("synthetic witness failed")
    return w
