"""Host G1 scalar multiplication (zk_g1_mul_host) against the oracle's
affine double-and-add, including the identity, k = 0, k = r - 1 and
scalars at or above r (reduced first)."""
import random

from oracle import curve as OC
from pyzkp_amd.curve import G1_GENERATOR, G1_ZERO, from_affine, g1_mul_native, normalize
from pyzkp_amd.field import FR_MODULUS


def test_g1_mul_host_matches_oracle():
    rng = random.Random(0x61)
    p_aff = OC.g1_mul_affine(OC.G1_GENERATOR, 0x1234567)
    ks = [0, 1, 2, 15, 16, 17, FR_MODULUS - 1, FR_MODULUS, FR_MODULUS + 3] + [rng.getrandbits(256) for _ in range(12)]
    for k in ks:
        got = normalize(g1_mul_native(from_affine(p_aff), k))
        assert got == OC.g1_mul_affine(p_aff, k % FR_MODULUS), k


def test_g1_mul_host_identity_and_order():
    assert normalize(g1_mul_native(G1_ZERO, 12345)) is None
    assert normalize(g1_mul_native(G1_GENERATOR, FR_MODULUS)) is None
    neg = g1_mul_native(G1_GENERATOR, FR_MODULUS - 1)
    g = normalize(G1_GENERATOR)
    assert normalize(neg) == (g[0], (-g[1]) % OC.Q)
