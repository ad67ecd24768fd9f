"""nise_2017_amd: MI355X (gfx950) implementation of NISE's wavefunction
propagation hot path.

The product is the C-ABI shared library nise_2017_amd/lib/libnise_hip.so
(include/nise_abi.h).  This package holds its sources (csrc/), the build
script, a ctypes view for tests and benchmarks (abi.py) and the synthetic
trajectory generator used by the benchmark (synth.py).
"""
from .abi import lib, TNon, make_non  # noqa: F401

__all__ = ["lib", "TNon", "make_non"]
