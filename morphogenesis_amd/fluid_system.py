"""Python host mirror of the reference's FluidSystem step API over libmg_sph.so.

Method names follow src/fluid_system.h:590-651 (Run2PhysicalSort, Run2InnerPhysicalLoop,
Run2GeneAction, InsertParticlesCL, PrefixSumCellsCL, CountingSortFullCL, ComputePressureCL,
ComputeForceCL, AdvanceCL, ComputeDiffusionCL, ComputeGenesCL, ZeroVelCL, TransferToCL,
TransferFromCL, setFreeze). Errors raise MgError instead of calling exit()
(src/fluid_system.cpp:447-452). The C++ facade in csrc/fluid_system.hpp offers the same
interface to C++ callers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi as F


class FluidSystem:
    """One simulation instance on the GPU (a handle of libmg_sph.so)."""

    prefix = "mg"

    def __init__(self, max_points: int, bonds_per_particle: int, lib=None):
        self.lib = lib or F.load_library()
        self.n = int(max_points)
        self.b = int(bonds_per_particle)
        self.c = 0
        self.h = C.c_void_p()
        self._call("create", C.byref(self.h), self.n, self.b)
        self.m_Time = 0.0
        self.params = self.grid = self.genome = None
        self.sim_params = None

    # ------------------------------------------------------------------ plumbing
    def _call(self, name, *args):
        r = getattr(self.lib, f"{self.prefix}_{name}")(*args)
        if r != F.MG_OK:
            raise F.MgError(f"{self.prefix}_{name} -> {F.ERRORS.get(r, r)}: {self.last_error()}")
        return r

    def last_error(self) -> str:
        return self.lib.mg_last_error().decode(errors="replace")

    def close(self):
        if self.h:
            self.lib.mg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ set-up
    def set_grid(self, grid: F.MgGrid):
        self._call("set_grid", self.h, C.byref(grid))
        self.c = int(grid.total)
        self.grid = grid

    def set_params(self, params: F.MgParams):
        self._call("set_params", self.h, C.byref(params))
        self.params = params

    def set_genome(self, genome: F.MgGenome):
        self._call("set_genome", self.h, C.byref(genome))
        self.genome = genome

    def setFreeze(self, freeze: bool):
        self._call("set_freeze", self.h, int(bool(freeze)))

    def set_sph_kernels(self, mode: int):
        """F.SPH_REFERENCE (the reference's Wendland-C2 density and (sr - r)^3 force, the
        default) or F.SPH_POLY6_SPIKY (Mueller 2003 poly6 density, spiky pressure and
        viscosity-Laplacian force)."""
        self._call("set_sph_kernels", self.h, int(mode))

    def upload(self, buf: int, arr: np.ndarray):
        shape, dtype = F.buffer_shape(buf, self.n, self.c, self.b)
        a = np.ascontiguousarray(arr, dtype=dtype).reshape(shape)
        self._call("upload", self.h, int(buf), a.ctypes.data_as(C.c_void_p), a.nbytes)

    def download(self, buf: int) -> np.ndarray:
        shape, dtype = F.buffer_shape(buf, self.n, self.c, self.b)
        a = np.empty(shape, dtype)
        self._call("download", self.h, int(buf), a.ctypes.data_as(C.c_void_p), a.nbytes)
        return a

    def TransferToCL(self, buffers: dict):
        for k, v in buffers.items():
            self.upload(k, v)

    def TransferFromCL(self, which=None) -> dict:
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
        self._call("sync", self.h)

    # ------------------------------------------------------------------ phases
    def InsertParticlesCL(self, readback: bool = False):
        if not readback:
            self._call("insert_particles", self.h, None, None, None)
            return None
        gcell = np.empty(self.n, np.uint32)
        gndx = np.empty(self.n, np.uint32)
        gcnt = np.empty(self.c, np.uint32)
        P = C.POINTER(C.c_uint32)
        self._call("insert_particles", self.h, gcell.ctypes.data_as(P), gndx.ctypes.data_as(P),
                   gcnt.ctypes.data_as(P))
        return gcell, gndx, gcnt

    def PrefixSumCellsCL(self, zero_offsets: int = 1):
        self._call("prefix_sum_cells", self.h, int(zero_offsets))

    def CountingSortFullCL(self):
        self._call("counting_sort_full", self.h)

    def ComputePressureCL(self):
        self._call("compute_pressure", self.h)

    def ComputeForceCL(self):
        self._call("compute_force", self.h)

    def AdvanceCL(self, time: float, dt: float, ss: float = 1.0):
        self._call("advance", self.h, float(time), float(dt), float(ss))

    def ComputeDiffusionCL(self):
        self._call("compute_diffusion", self.h)

    def ComputeGenesCL(self):
        self._call("compute_genes", self.h)

    def InitializeBondsCL(self):
        """initialize_bonds over gene 1's dense list (src/host_CUDA.cpp:650-660); needs the
        sorted state."""
        self._call("initialize_bonds", self.h)

    # ------------------------------------------------------------------ remodelling
    def ComputeBondChangesCL(self, steps_per_inner_loop: int):
        """computeBondChanges over gene 2's dense list (src/host_CL.cpp:2057-2086): remodels
        rest length and modulus from the averaged strain integrators and counts each particle's
        first broken or empty bond slot for heal."""
        self._call("compute_bond_changes", self.h, int(steps_per_inner_loop))

    def PrefixSumChangesCL(self, zero_offsets: int = 1):
        """Per-cell offsets and lengths of the nine change lists (src/host_CL.cpp:865-1062)."""
        self._call("prefix_sum_changes", self.h, int(zero_offsets))

    def CountingSortChangesCL(self):
        """The (slot, bond) pairs of each change list (src/host_CL.cpp:1415-1519)."""
        self._call("counting_sort_changes", self.h)

    def ComputeParticleChangesCL(self, seed: int | None = None):
        """heal over list 0 (src/host_CL.cpp:2088-2173). The random shifts come from
        (seed, particle ID); without a seed the handle's remodelling counter is used."""
        if seed is None:
            seed = self._heal_seed
            self._heal_seed += 1
        self._call("compute_particle_changes", self.h, int(seed) & 0xFFFFFFFFFFFFFFFF)

    _heal_seed = 0

    def Run2Remodelling(self, steps_per_inner_loop: int, seed: int | None = None):
        """src/fluid_system.cpp:2037-2071: bond changes, their lists, then heal."""
        self.ComputeBondChangesCL(steps_per_inner_loop)
        self.PrefixSumChangesCL(1)
        self.CountingSortChangesCL()
        self.ComputeParticleChangesCL(seed)

    def ZeroVelCL(self):
        self._call("zero_vel", self.h)

    def Run2PhysicalSort(self):
        self._call("physical_sort", self.h)

    def Run2InnerPhysicalLoop(self, dt: float):
        self._call("inner_physical_loop", self.h, float(self.m_Time), float(dt))
        self.m_Time += dt

    def Run2GeneAction(self):
        self._call("gene_action", self.h)

    def step_learned(self, net, nsteps: int, dt: float, flags: int = 0):
        """nsteps learned coarse steps of dt (mg_step_learned): the step's pressure and force
        passes, the network, and the advance with the net's drift-velocity correction."""
        from .learned import _declare

        _declare(self.lib)
        self._call("step_learned", self.h, net.h, int(nsteps), float(dt), int(flags))
        self.m_Time += nsteps * dt

    def learned_fallbacks(self) -> int:
        from .learned import _declare

        _declare(self.lib)
        c = C.c_uint64()
        self._call("learned_fallbacks", self.h, C.byref(c))
        return int(c.value)

    def step(self, nsteps: int, dt: float, flags: int = 0):
        self._call("step", self.h, int(nsteps), float(dt), int(flags))
        self.m_Time += nsteps * dt

    # ------------------------------------------------------------------ reference files
    @classmethod
    def from_reference_files(cls, sim_params_path: str, points_path: str,
                             genome_path: str | None = None, lib=None,
                             check_bounds: bool = True) -> "FluidSystem":
        """ReadSimParams + ReadGenome + ReadPointsCSV2 (src/fileCL_IO.cpp:71-157,336-530):
        a system sized to the point file (mMaxPoints = its rows, B from its header). With
        check_bounds, a particle outside PBOUNDMIN/MAX or faster than PVEL_LIMIT is an
        error, as in the reference's reader."""
        from . import fileio

        sc = fileio.load_scene(sim_params_path, points_path, genome_path,
                               check_bounds=check_bounds)
        sim = cls(sc.n, sc.b, lib=lib)
        sim.load_scene(sc)
        sim.m_Time = sc.info["time"]
        sim.m_DT = sc.dt
        sim.sim_params = sc.info["sim_params"]
        return sim

    def SavePointsCSV2(self, directory: str, frame: int) -> str:
        from . import fileio

        return fileio.save_points_csv2(self.TransferFromCL(), directory, frame)

    def SavePointsVTP2(self, directory: str, frame: int) -> str:
        """SavePointsVTP2 (src/fileCL_IO.cpp:1223-1684): VTK XML PolyData of the state."""
        from . import fileio

        g = self.grid
        vmin = g.gmin[:] if g is not None else (0.0, 0.0, 0.0)
        vmax = g.gmax[:] if g is not None else (0.0, 0.0, 0.0)
        return fileio.save_points_vtp(self.TransferFromCL(), directory, frame, vmin, vmax)

    def WriteGenome(self, directory: str) -> str:
        from . import fileio

        if self.genome is None:
            raise F.MgError("WriteGenome: no genome set")
        return fileio.write_genome(self.genome, directory)

    def WriteSimParams(self, directory: str) -> str:
        from . import fileio

        if self.sim_params is None:
            raise F.MgError("WriteSimParams: the system was not set up from SimParams")
        sp = self.sim_params.copy()
        sp.time = self.m_Time
        return fileio.write_sim_params(sp, directory)

    def save_checkpoint(self, path: str) -> str:
        from . import fileio

        return fileio.save_checkpoint(self, path, genome=self.genome, params=self.params,
                                      grid=self.grid)

    @classmethod
    def from_checkpoint(cls, path: str, lib=None) -> "FluidSystem":
        """A system in the state save_checkpoint wrote; stepping it continues the saved run
        bit for bit."""
        from . import fileio

        n, b, bufs, params, grid, genome, time = fileio.load_checkpoint(path)
        if params is None or grid is None or genome is None:
            raise F.MgError(f"{path}: checkpoint lacks parameters, grid or genome")
        sim = cls(n, b, lib=lib)
        sim.set_grid(grid)
        sim.set_params(params)
        sim.set_genome(genome)
        sim.TransferToCL(bufs)
        # the reference's run loop starts with a physical sort; on a sorted state it moves
        # nothing and restores the point counts
        sim.Run2PhysicalSort()
        sim.m_Time = time
        return sim
