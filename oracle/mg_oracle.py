"""Python wrapper of the CPU oracle (liboracle built from oracle/mg_oracle.cpp).

TEST INFRASTRUCTURE: used only by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline, always as the checker or the reported baseline, never as the product path.
The interface matches morphogenesis_amd.fluid_system.FluidSystem so both can be driven by
the same test code.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from morphogenesis_amd import _ffi as F

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmg_oracle.so")

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing; run __graft_entry__.build() first")
        _lib = F.declare(C.CDLL(LIB_PATH), "mgo")
        _lib.mgo_acosf.argtypes = [C.c_float]
        _lib.mgo_acosf.restype = C.c_float
        for name in ("mgo_tanhf", "mgo_expf"):
            getattr(_lib, name).argtypes = [C.c_float]
            getattr(_lib, name).restype = C.c_float
        _lib.mgo_heal_word.argtypes = [C.c_uint64, C.c_uint32]
        _lib.mgo_heal_word.restype = C.c_uint32
    return _lib


def acosf(x: float) -> float:
    """The float acos the bond-angle score uses (Cephes method, shared with the kernels)."""
    return float(load().mgo_acosf(float(x)))


def tanhf(x: float) -> float:
    """The float tanh of the bond remodelling (Cephes method, shared with the kernels)."""
    return float(load().mgo_tanhf(float(x)))


def expf(x: float) -> float:
    """The float exp inside tanhf."""
    return float(load().mgo_expf(float(x)))


def heal_word(seed: int, pid: int) -> int:
    """heal's 32-bit random word for a particle ID (SplitMix64 output function)."""
    return int(load().mgo_heal_word(int(seed) & 0xFFFFFFFFFFFFFFFF, int(pid)))


class Oracle:
    prefix = "mgo"

    def __init__(self, max_points: int, bonds_per_particle: int, precision: int = 32,
                 threads: int = 0):
        self.lib = load()
        self.n = int(max_points)
        self.b = int(bonds_per_particle)
        self.c = 0
        self.h = C.c_void_p()
        self._call("create", C.byref(self.h), self.n, self.b, int(precision))
        if threads:
            self._call("set_threads", self.h, int(threads))
        self.m_Time = 0.0

    def _call(self, name, *args):
        r = getattr(self.lib, f"mgo_{name}")(*args)
        if r != F.MG_OK:
            raise F.MgError(f"mgo_{name} -> {F.ERRORS.get(r, r)}")
        return r

    def close(self):
        if self.h:
            self.lib.mgo_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grid(self, grid):
        self._call("set_grid", self.h, C.byref(grid))
        self.c = int(grid.total)

    def set_params(self, params):
        self._call("set_params", self.h, C.byref(params))

    def set_genome(self, genome):
        self._call("set_genome", self.h, C.byref(genome))

    def setFreeze(self, freeze: bool):
        self._call("set_freeze", self.h, int(bool(freeze)))

    def set_sph_kernels(self, mode: int):
        self._call("set_sph_kernels", self.h, int(mode))

    def upload(self, buf, arr):
        shape, dtype = F.buffer_shape(buf, self.n, self.c, self.b)
        a = np.ascontiguousarray(arr, dtype=dtype).reshape(shape)
        self._call("upload", self.h, int(buf), a.ctypes.data_as(C.c_void_p), a.nbytes)

    def download(self, buf):
        shape, dtype = F.buffer_shape(buf, self.n, self.c, self.b)
        a = np.empty(shape, dtype)
        self._call("download", self.h, int(buf), a.ctypes.data_as(C.c_void_p), a.nbytes)
        return a

    def TransferToCL(self, buffers):
        for k, v in buffers.items():
            self.upload(k, v)

    def TransferFromCL(self, which=None):
        ids = which if which is not None else F.PARTICLE_BUFFERS
        return {k: self.download(k) for k in ids}

    def load_scene(self, scene):
        self.set_grid(scene.grid)
        self.set_params(scene.params)
        self.set_genome(scene.genome)
        self.TransferToCL(scene.buffers)

    def counts(self):
        a, b = C.c_uint32(), C.c_uint32()
        self._call("get_counts", self.h, C.byref(a), C.byref(b))
        return int(a.value), int(b.value)

    def sync(self):
        pass

    def InsertParticlesCL(self, readback: bool = False):
        self._call("insert_particles", self.h)
        if readback:
            return self.download(F.FGCELL), self.download(F.FGNDX), self.download(F.FBIN_COUNT)
        return None

    def PrefixSumCellsCL(self, zero_offsets: int = 1):
        self._call("prefix_sum_cells", self.h)

    def CountingSortFullCL(self):
        self._call("counting_sort_full", self.h)

    def ComputePressureCL(self):
        self._call("compute_pressure", self.h)

    def ComputeForceCL(self):
        self._call("compute_force", self.h)

    def AdvanceCL(self, time, dt, ss=1.0):
        self._call("advance", self.h, float(time), float(dt), float(ss))

    def ComputeDiffusionCL(self):
        self._call("compute_diffusion", self.h)

    def ComputeGenesCL(self):
        self._call("compute_genes", self.h)

    def InitializeBondsCL(self):
        self._call("initialize_bonds", self.h)

    def ComputeBondChangesCL(self, steps_per_inner_loop):
        self._call("compute_bond_changes", self.h, int(steps_per_inner_loop))

    def PrefixSumChangesCL(self, zero_offsets=1):
        self._call("prefix_sum_changes", self.h, int(zero_offsets))

    def CountingSortChangesCL(self):
        self._call("counting_sort_changes", self.h)

    def ComputeParticleChangesCL(self, seed):
        self._call("compute_particle_changes", self.h, int(seed) & 0xFFFFFFFFFFFFFFFF)

    def Run2Remodelling(self, steps_per_inner_loop, seed):
        self.ComputeBondChangesCL(steps_per_inner_loop)
        self.PrefixSumChangesCL(1)
        self.CountingSortChangesCL()
        self.ComputeParticleChangesCL(seed)

    def ZeroVelCL(self):
        self._call("zero_vel", self.h)

    def Run2PhysicalSort(self):
        self._call("physical_sort", self.h)

    def Run2InnerPhysicalLoop(self, dt):
        self._call("inner_physical_loop", self.h, float(self.m_Time), float(dt))
        self.m_Time += dt

    def Run2GeneAction(self):
        self._call("gene_action", self.h)

    def step(self, nsteps, dt, flags=0):
        self._call("step", self.h, int(nsteps), float(dt), int(flags))
        self.m_Time += nsteps * dt
