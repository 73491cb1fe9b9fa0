"""The R1CS builder and the reference's example circuits on the CPU: the
witness satisfies its R1CS and a wrong public value does not (the
reference's own checks, tests/correctness/cpu/test_poseidon.py and
test_groth16_e2e.py); the Poseidon circuit equals the direct permutation;
and the oracle's Groth16 over the cubic circuit satisfies the pairing
equation in the exponent."""
import pytest

from oracle import groth16 as OG
from pyzkp_amd.circuits import check_r1cs, cubic_circuit, poseidon_circuit, poseidon_hash
from pyzkp_amd.field import FR_MODULUS, words_to_ints


def test_cubic_circuit():
    circ, w = cubic_circuit(9)
    vals = words_to_ints(w)
    assert vals[circ.public_ids[1]] == (9 ** 3 + 9 + 5) % FR_MODULUS
    assert circ.n_constraints == 3
    check_r1cs(circ, w)
    bad, wb = cubic_circuit(9, y=(9 ** 3 + 9 + 6))
    with pytest.raises(AssertionError):
        check_r1cs(bad, wb)


def test_poseidon_circuit():
    h = poseidon_hash([12345, 67890])
    circ, w = poseidon_circuit(12345, 67890)
    assert words_to_ints(w)[circ.public_ids[1]] == h
    # full rounds: 3 x 3 products, partial rounds: 3 products, plus the
    # assertion; the capacity element is still a constant in the first
    # round, so its S-box costs no constraint
    assert circ.n_constraints == 8 * 9 + 53 * 3 + 1 - 3
    check_r1cs(circ, w)
    bad, wb = poseidon_circuit(12345, 67890, expected=(h + 1) % FR_MODULUS)
    with pytest.raises(AssertionError):
        check_r1cs(bad, wb)
    assert poseidon_hash([12345, 67891]) != h


def test_oracle_groth16_over_cubic():
    circ, w = cubic_circuit(9)
    wi = words_to_ints(w)
    A, B, C = circ.A.row_dicts(), circ.B.row_dicts(), circ.C.row_dicts()
    td = (1234567, 89, 1011, 1213, 1415)
    opk = OG.setup(A, B, C, circ.n, circ.n_vars, circ.public_ids, list(circ.aux_ids), td)
    pf = OG.prove(opk, A, B, C, wi, 5, 7)
    assert OG.check_in_exponent(opk, wi, 5, 7, pf, circ.n)
