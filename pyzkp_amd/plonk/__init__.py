"""PLONK prover pieces on the HIP path."""
from .quotient import compute_quotient, draw_shift

__all__ = ["compute_quotient", "draw_shift"]
