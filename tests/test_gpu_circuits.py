"""The reference's example circuits end to end through the HIP Groth16
path (tests/correctness/cpu/test_groth16_e2e.py, test_poseidon.py): the
proof verifies, a wrong public value fails verification, a witness for a
wrong claim is refused by the prover, and proving key and proof equal the
oracle's point for point."""
import pytest

from oracle import groth16 as OG
from pyzkp_amd.circuits import cubic_circuit, poseidon_circuit, poseidon_hash
from pyzkp_amd.curve import normalize
from pyzkp_amd.field import FR_MODULUS, words_to_ints
from pyzkp_amd.g2 import normalize as normalize2
from pyzkp_amd.groth16 import prove, setup, verify

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["cubic", "poseidon"])
def test_circuit_end_to_end(zctx, name):
    if name == "cubic":
        circ, w = cubic_circuit(9)
        wrong, wb = cubic_circuit(9, y=9 ** 3 + 9 + 6)
    else:
        circ, w = poseidon_circuit(12345, 67890)
        wrong, wb = poseidon_circuit(12345, 67890, expected=(poseidon_hash([12345, 67890]) + 1) % FR_MODULUS)
    pk = setup(circ, zctx, seed=0x6100)
    pub = words_to_ints(w[circ.public_ids])
    pf = prove(pk, w, 31337, 4242)
    assert verify(pk.vk, pub, pf)
    assert not verify(pk.vk, [pub[0], (pub[1] + 1) % FR_MODULUS], pf)
    with pytest.raises(ValueError):
        prove(pk, wb, 1, 2)   # same circuit shape, the witness violates the assertion row
    wi = words_to_ints(w)
    A, B, C = circ.A.row_dicts(), circ.B.row_dicts(), circ.C.row_dicts()
    td = (pk.trapdoor_tau, pk.trapdoor_alpha, pk.trapdoor_beta, pk.trapdoor_gamma, pk.trapdoor_delta)
    opk = OG.setup(A, B, C, circ.n, circ.n_vars, circ.public_ids, list(circ.aux_ids), td)
    opf = OG.prove(opk, A, B, C, wi, 31337, 4242)
    assert normalize(pf.a) == opf["a"] and normalize2(pf.b) == opf["b"] and normalize(pf.c) == opf["c"]
    pk.free()
