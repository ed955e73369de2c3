"""MI355X-native Rayleigh-Chebyshev lowest-eigenpair solver.

librcg.so (HIP, gfx950) holds the kernels and the solve; ``solver`` mirrors
the reference's RayleighChebyshevLM interface over its C ABI, ``operators``
builds the synthetic operators and the start-vector stream.
"""
from . import operators  # noqa: F401

__all__ = ["operators", "capi", "solver"]
