"""Groth16 over the HIP path: synthetic circuits, device setup, device
prover (G2 MSM on host threads), pairing verifier."""
from .prove import prove, prove_many, release_work
from .setup import setup
from .synthetic import CSR, SyntheticCircuit, synthetic_circuit, synthetic_witness
from .types import Proof, ProvingKey, VerifyingKey
from .verify import verify

__all__ = ["prove", "prove_many", "release_work", "setup", "CSR", "SyntheticCircuit", "synthetic_circuit",
           "synthetic_witness", "Proof", "ProvingKey", "VerifyingKey", "verify"]
