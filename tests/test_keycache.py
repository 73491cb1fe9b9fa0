"""Proving-key cache: the fingerprint follows the circuit, and a key file
written for another circuit, scheme, version or format is refused with the
reference's CacheMismatchError messages (runtime/cache.py:85-97) before
anything touches the device; the file loads without executing code."""
import json

import numpy as np
import pytest

from pyzkp_amd.groth16 import synthetic_circuit
from pyzkp_amd.groth16.keycache import CacheMismatchError, circuit_fingerprint, default_key_path, load_key


def test_fingerprint_tracks_the_circuit():
    c1 = synthetic_circuit(6, 0x10)
    assert circuit_fingerprint(c1) == circuit_fingerprint(synthetic_circuit(6, 0x10))
    assert circuit_fingerprint(c1) != circuit_fingerprint(synthetic_circuit(6, 0x11))   # other mix constants
    assert circuit_fingerprint(c1) != circuit_fingerprint(synthetic_circuit(7, 0x10))
    c2 = synthetic_circuit(6, 0x10)
    c2.A.coeffs = c2.A.coeffs.copy()
    c2.A.coeffs[0, 0] ^= 1
    assert circuit_fingerprint(c2) != circuit_fingerprint(c1)
    assert len(circuit_fingerprint(c1)) == 16
    assert circuit_fingerprint(c1) in default_key_path("/tmp/x", c1)


@pytest.mark.parametrize("field,value,msg", [("format", "other", "format"), ("version", 99, "version"),
                                             ("scheme", "plonk", "scheme"), ("field_modulus", "7", "modulus"),
                                             ("ir_fingerprint", "0" * 16, "circuit")])
def test_meta_mismatch(tmp_path, field, value, msg):
    circ = synthetic_circuit(6, 0x10)
    meta = {"format": "setup_cache", "version": 1, "scheme": "groth16",
            "field_modulus": "21888242871839275222246405745257275088548364400416034343698204186575808495617",
            "ir_fingerprint": circuit_fingerprint(circ)}
    meta[field] = value
    path = tmp_path / "k.npz"
    np.savez(path, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    with pytest.raises(CacheMismatchError, match=msg):
        load_key(str(path), circ, ctx=None)
