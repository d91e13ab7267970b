"""morphogenesis_amd — MI355X-native (gfx950) per-timestep SPH step of Morphogenesis.

The product is libmg_sph.so (hand-written HIP kernels behind the C-ABI of include/mg_sph.h)
plus its host facades: the C++ class mg::FluidSystem (csrc/fluid_system.hpp) and the Python
mirror morphogenesis_amd.fluid_system.FluidSystem. Scene generators for the benchmark
workloads live in morphogenesis_amd.scenes.
"""
from . import _ffi
from ._ffi import MgError, MgGenome, MgGrid, MgParams, load_library

__all__ = ["_ffi", "MgError", "MgGenome", "MgGrid", "MgParams", "load_library"]
