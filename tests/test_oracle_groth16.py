"""Pins the Groth16 oracle: its proofs satisfy the verification equation in
the exponent (independent derivation from the trapdoor).  CPU only."""
import pytest

from oracle import groth16 as OG
from oracle.field import R_MOD


def _tiny():
    from pyzkp_amd.groth16 import synthetic_circuit, synthetic_witness
    from pyzkp_amd.field import words_to_ints
    circ = synthetic_circuit(5, 0x51)
    w = words_to_ints(synthetic_witness(circ, 0x5101))
    return circ, w


def test_oracle_groth16_exponent_check():
    circ, w = _tiny()
    A, B, C = circ.A.row_dicts(), circ.B.row_dicts(), circ.C.row_dicts()
    td = (1234567, 7654321, 2222223, 3333337, 4444447)
    pk = OG.setup(A, B, C, circ.n, circ.n_vars, circ.public_ids, list(circ.aux_ids), td)
    pf = OG.prove(pk, A, B, C, w, 11111, 22222)
    assert OG.check_in_exponent(pk, w, 11111, 22222, pf, circ.n)
    # h (x) (X^n - 1) == A B - C as polynomials
    assert len(pf["h"]) == circ.n - 1
    bad = list(w)
    bad[0] = (bad[0] + 1) % R_MOD
    with pytest.raises(ValueError):
        OG.prove(pk, A, B, C, bad, 1, 2)
    # tampered proof fails the check
    pf2 = dict(pf)
    pf2["c"] = OG.C1.aff_add(pf["c"], OG.C1.G1_GENERATOR)
    assert not OG.check_in_exponent(pk, w, 11111, 22222, pf2, circ.n)


def test_exponent_key_equals_setup_columns():
    """The Lagrange-at-tau route (exponent_key) and setup's iNTT-then-Horner
    route give the same QAP columns at tau, and the proof checks against
    either key."""
    circ, w = _tiny()
    A, B, C = circ.A.row_dicts(), circ.B.row_dicts(), circ.C.row_dicts()
    td = (98765, 4321, 55555, 77777, 99991)
    full = OG.setup(A, B, C, circ.n, circ.n_vars, circ.public_ids, list(circ.aux_ids), td)
    ek = OG.exponent_key(A, B, C, circ.n, circ.n_vars, circ.public_ids, td)
    for f in ("a_tau", "b_tau", "k_tau"):
        assert ek[f] == full[f], f
    pf = OG.prove(full, A, B, C, w, 5, 6)
    assert OG.check_in_exponent(ek, w, 5, 6, pf, circ.n)
    assert not OG.check_in_exponent(ek, w, 5, 7, pf, circ.n)
