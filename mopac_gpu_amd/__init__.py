"""MI355X (gfx950) implementation of MOPAC's SCF inner loop.

The compute path is libmopachip.so (hand-written HIP kernels behind the C ABI
in include/mopac_hip.h); this package holds its Python mirror of the
reference routines and the input generators.
"""
from . import lib  # noqa: F401

__all__ = ["lib", "geometry", "hamiltonian", "scf"]
