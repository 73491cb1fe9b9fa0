"""Host pairing: bilinearity, non-degeneracy and the reference key's SRS
(e(tau G1, G2) == e(G1, tau G2)).  CPU only."""
import json
import os
import random

import pytest

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pyzkp_amd", "libzkhip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libzkhip.so not built")

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def test_bilinear_and_nondegenerate():
    from pyzkp_amd.curve import G1_GENERATOR, g1_mul, g1_neg
    from pyzkp_amd.g2 import G2_GENERATOR, g2_mul
    from pyzkp_amd.pairing import pairing, pairing_check
    one = (1,) + (0,) * 11
    e = pairing(G1_GENERATOR, G2_GENERATOR)
    assert e != one
    rng = random.Random(3)
    a, b = rng.randrange(1, R), rng.randrange(1, R)
    assert pairing(g1_mul(G1_GENERATOR, a), g2_mul(G2_GENERATOR, b)) == pairing(g1_mul(G1_GENERATOR, a * b % R), G2_GENERATOR)
    assert pairing(g1_mul(G1_GENERATOR, a), G2_GENERATOR) == pairing(G1_GENERATOR, g2_mul(G2_GENERATOR, a))
    assert pairing_check([(g1_mul(G1_GENERATOR, a), G2_GENERATOR), (g1_neg(G1_GENERATOR), g2_mul(G2_GENERATOR, a))])
    assert not pairing_check([(g1_mul(G1_GENERATOR, a), G2_GENERATOR), (g1_neg(G1_GENERATOR), g2_mul(G2_GENERATOR, a + 1))])
    # identity pairs contribute 1
    assert pairing_check([((1, 1, 0), G2_GENERATOR)])


def test_reference_srs_pairing():
    from pyzkp_amd.pairing import pairing
    with open(os.path.join(os.path.dirname(__file__), "golden", "plonk_pk_n16.json")) as f:
        pk = json.load(f)
    g1 = [(p[0], p[1], 1) for p in pk["srs_g1"][:3]]
    g2 = [(tuple(p[0]), tuple(p[1]), (1, 0)) for p in pk["srs_g2"]]
    # e(tau G1, G2) == e(G1, tau G2)
    assert pairing(g1[1], g2[0]) == pairing(g1[0], g2[1])
    # e(tau^2 G1, G2) == e(tau G1, tau G2)
    assert pairing(g1[2], g2[0]) == pairing(g1[1], g2[1])
