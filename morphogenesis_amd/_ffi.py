"""ctypes mirror of include/mg_types.h and loader for libmg_sph.so.

The structures below must match include/mg_types.h field for field; a size check
against the C library (mg_sizeof_*) guards against drift.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libmg_sph.so")

# buffer ids (src/fluid.h:138-206)
FPOS, FVEL, FVEVAL, FFORCE, FPRESS, FDENSITY, FAGE, FCOLOR, FGCELL, FGNDX = range(10)
FELASTIDX, FNERVEIDX, FCONC, FEPIGEN, FBIN, FBIN_COUNT, FBIN_OFFSET = 14, 15, 16, 17, 18, 19, 20
FPARTICLEIDX, FPARTICLE_ID, FMASS_RADIUS = 29, 30, 31
FBIN_COUNT_ACTIVE_GENES, FBIN_OFFSET_ACTIVE_GENES, FDENSE_LIST_LENGTHS, FDENSE_LISTS = 32, 33, 34, 35
# remodelling change lists (src/fluid.h:197-201)
FBIN_OFFSET_CHANGES, FBIN_COUNT_CHANGES, FDENSE_LIST_LENGTHS_CHANGES, FDENSE_LISTS_CHANGES = 37, 38, 39, 40

PARTICLE_BUFFERS = [FPOS, FVEL, FVEVAL, FFORCE, FPRESS, FDENSITY, FAGE, FCOLOR, FGCELL, FGNDX,
                    FELASTIDX, FNERVEIDX, FCONC, FEPIGEN, FBIN, FPARTICLEIDX, FPARTICLE_ID,
                    FMASS_RADIUS]
GRID_BUFFERS = [FBIN_COUNT, FBIN_OFFSET, FBIN_COUNT_ACTIVE_GENES, FBIN_OFFSET_ACTIVE_GENES,
                FDENSE_LIST_LENGTHS, FDENSE_LISTS]
BUFFER_NAMES = {
    FPOS: "FPOS", FVEL: "FVEL", FVEVAL: "FVEVAL", FFORCE: "FFORCE", FPRESS: "FPRESS",
    FDENSITY: "FDENSITY", FAGE: "FAGE", FCOLOR: "FCOLOR", FGCELL: "FGCELL", FGNDX: "FGNDX",
    FELASTIDX: "FELASTIDX", FNERVEIDX: "FNERVEIDX", FCONC: "FCONC", FEPIGEN: "FEPIGEN",
    FBIN: "FBIN", FBIN_COUNT: "FBIN_COUNT", FBIN_OFFSET: "FBIN_OFFSET",
    FPARTICLEIDX: "FPARTICLEIDX", FPARTICLE_ID: "FPARTICLE_ID", FMASS_RADIUS: "FMASS_RADIUS",
    FBIN_COUNT_ACTIVE_GENES: "FBIN_COUNT_ACTIVE_GENES",
    FBIN_OFFSET_ACTIVE_GENES: "FBIN_OFFSET_ACTIVE_GENES",
    FDENSE_LIST_LENGTHS: "FDENSE_LIST_LENGTHS", FDENSE_LISTS: "FDENSE_LISTS",
    FBIN_OFFSET_CHANGES: "FBIN_OFFSET_CHANGES", FBIN_COUNT_CHANGES: "FBIN_COUNT_CHANGES",
    FDENSE_LIST_LENGTHS_CHANGES: "FDENSE_LIST_LENGTHS_CHANGES",
    FDENSE_LISTS_CHANGES: "FDENSE_LISTS_CHANGES",
}
CHANGE_BUFFERS = [FBIN_OFFSET_CHANGES, FBIN_COUNT_CHANGES, FDENSE_LIST_LENGTHS_CHANGES,
                  FDENSE_LISTS_CHANGES]

NUM_TF = 16
NUM_GENES = 16
NUM_CHANGES = 9
DATA_PER_BOND = 9
NONE = 0xFFFFFFFF
GRID_UNDEF = 0xFFFFFFFF
STEP_GENES = 1
STEP_GRAPH = 2
STEP_SPECIAL = 4
# SPH kernels (mg_set_sph_kernels): the reference's live ones, or Mueller 2003 poly6/spiky
SPH_REFERENCE = 0
SPH_POLY6_SPIKY = 1

MG_OK = 0
ERRORS = {-1: "MG_ERR_ARG", -2: "MG_ERR_HIP", -3: "MG_ERR_STATE", -4: "MG_ERR_UNSUPPORTED",
          -5: "MG_ERR_SIZE", -6: "MG_ERR_ALLOC"}


class MgGenome(C.Structure):
    _fields_ = [
        ("mutability", C.c_uint32 * NUM_GENES),
        ("delay", C.c_uint32 * NUM_GENES),
        ("sensitivity", (C.c_uint32 * NUM_GENES) * NUM_GENES),
        ("tf_diffusability", C.c_uint32 * NUM_TF),
        ("tf_breakdown_rate", C.c_uint32 * NUM_TF),
        ("secrete", (C.c_int32 * (2 * NUM_TF + 1)) * NUM_GENES),
        ("activate", (C.c_int32 * (2 * NUM_GENES + 1)) * NUM_GENES),
        ("tanh_param", (C.c_float * 8) * 3),
        ("param", (C.c_float * 12) * 3),
    ]


class MgParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "psimscale", "psmoothradius", "pradius", "r2", "pmass", "prest_dens",
        "pextstiff", "pintstiff", "pvisc", "psurface_t", "pdamp",
        "pforce_min", "pforce_max", "pforce_freq", "pground_slope")] + [
        ("pgravity", C.c_float * 3), ("pboundmin", C.c_float * 3), ("pboundmax", C.c_float * 3),
    ] + [(n, C.c_float) for n in (
        "AL", "AL2", "VL", "VL2", "pdist", "poly6kern", "spikykern", "lapkern", "gausskern",
        "wendlandC2kern", "H", "d2", "rd2", "vterm", "actuation_factor", "actuation_period",
        "diff_scale", "secrete_scale")] + [("freeze", C.c_int32), ("frame", C.c_uint32)]


class MgGrid(C.Structure):
    _fields_ = [
        ("res", C.c_int32 * 3), ("scan_max", C.c_int32 * 3),
        ("size", C.c_float * 3), ("delta", C.c_float * 3),
        ("gmin", C.c_float * 3), ("gmax", C.c_float * 3),
        ("total", C.c_int32), ("srch", C.c_int32), ("adj_cnt", C.c_int32), ("chk", C.c_int32),
        ("adj", C.c_int32 * 64),
    ]


def f32(x) -> float:
    """Round a Python float to the nearest float32 (host float arithmetic of the reference)."""
    return float(np.float32(x))


_libm = None


def powf(x, y) -> np.float32:
    """The C library's powf, the function the reference's host pow(float, float) calls
    (src/fluid_system.cpp:548, 2529-2534). numpy's float32 power can differ by an ulp."""
    global _libm
    if _libm is None:
        _libm = C.CDLL("libm.so.6")
        _libm.powf.restype = C.c_float
        _libm.powf.argtypes = [C.c_float, C.c_float]
    return np.float32(_libm.powf(float(np.float32(x)), float(np.float32(y))))


def params_from_fluid_param(ss, sr, pr, mass, rest, bmin, bmax, estiff, istiff, visc, surface_t,
                            damp, fmin, fmax, ffreq, gslope, gx, gy, gz, al, vl, a_f=0.0,
                            a_p=1.0) -> MgParams:
    """FluidParamCL (src/fluid_system.cpp:521-610): the MgParams that mg_fluid_param stores,
    derived by the library's own host code (mg_derive_params, no device work)."""
    lib = load_library()
    p = MgParams()
    fl = C.c_float * 3
    r = lib.mg_derive_params(C.byref(p), ss, sr, pr, mass, rest, fl(*bmin), fl(*bmax), estiff,
                             istiff, visc, surface_t, damp, fmin, fmax, ffreq, gslope, gx, gy, gz,
                             al, vl, a_f, a_p)
    if r != MG_OK:
        raise MgError(f"mg_derive_params -> {ERRORS.get(r, r)}: {lib.mg_last_error().decode()}")
    return p


def grid_setup(vmin, vmax, sim_scale, cell_size, smooth_radius) -> MgGrid:
    """SetupGrid (src/fluid_system.cpp:2073-2131) + FluidSetupCL (src/host_CL.cpp:39-106),
    derived by the library's own host code (mg_derive_grid, no device work)."""
    lib = load_library()
    g = MgGrid()
    fl = C.c_float * 3
    r = lib.mg_derive_grid(C.byref(g), fl(*vmin), fl(*vmax), sim_scale, cell_size, smooth_radius)
    if r != MG_OK:
        raise MgError(f"mg_derive_grid -> {ERRORS.get(r, r)}: {lib.mg_last_error().decode()}")
    return g


def buffer_shape(buf: int, n: int, c: int, b: int):
    """numpy (shape, dtype) of a buffer in the reference layout."""
    u, f = np.uint32, np.float32
    table = {
        FPOS: ((n, 4), f), FVEL: ((n, 4), f), FVEVAL: ((n, 4), f), FFORCE: ((n, 4), f),
        FPRESS: ((n,), f), FDENSITY: ((n,), f), FAGE: ((n,), u), FCOLOR: ((n,), u),
        FGCELL: ((n,), u), FGNDX: ((n,), u), FNERVEIDX: ((n,), u), FBIN: ((n,), u),
        FPARTICLE_ID: ((n,), u), FMASS_RADIUS: ((n,), u),
        FELASTIDX: ((n, b, DATA_PER_BOND), u), FPARTICLEIDX: ((n, b, 2), u),
        FCONC: ((n, NUM_TF), f), FEPIGEN: ((NUM_GENES, n), u),
        FBIN_COUNT: ((c,), u), FBIN_OFFSET: ((c,), u),
        FBIN_COUNT_ACTIVE_GENES: ((NUM_GENES, c), u), FBIN_OFFSET_ACTIVE_GENES: ((NUM_GENES, c), u),
        FDENSE_LIST_LENGTHS: ((NUM_GENES,), u), FDENSE_LISTS: ((NUM_GENES, n), u),
        FBIN_OFFSET_CHANGES: ((NUM_CHANGES, c), u), FBIN_COUNT_CHANGES: ((NUM_CHANGES, c), u),
        FDENSE_LIST_LENGTHS_CHANGES: ((NUM_CHANGES,), u),
        FDENSE_LISTS_CHANGES: ((NUM_CHANGES, 2, n), u),
    }
    return table[buf]


class MgError(RuntimeError):
    pass


_lib = None


def load_library(path: str | None = None):
    """Load libmg_sph.so (the HIP product path). Raises if it is missing: there is no
    CPU fallback."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise MgError(f"{p} is missing; run __graft_entry__.build() (hipcc, gfx950) first")
    lib = C.CDLL(p)
    declare(lib, "mg")
    lib.mg_last_error.restype = C.c_char_p
    lib.mg_last_error.argtypes = []
    if path is None:
        _lib = lib
    return lib


def declare(lib, prefix: str):
    """Declare the shared entry-point signatures for the mg_ (HIP) or mgo_ (oracle) library."""
    h = C.c_void_p
    P = C.POINTER
    sig = {
        "destroy": [h],
        "set_params": [h, P(MgParams)],
        "set_genome": [h, P(MgGenome)],
        "set_freeze": [h, C.c_int],
        "set_sph_kernels": [h, C.c_int],
        "upload": [h, C.c_int, C.c_void_p, C.c_size_t],
        "download": [h, C.c_int, C.c_void_p, C.c_size_t],
        "get_counts": [h, P(C.c_uint32), P(C.c_uint32)],
        "prefix_sum_cells": [h] + ([C.c_int] if prefix == "mg" else []),
        "counting_sort_full": [h],
        "compute_pressure": [h],
        "compute_force": [h],
        "advance": [h, C.c_float, C.c_float, C.c_float],
        "compute_diffusion": [h],
        "compute_genes": [h],
        "zero_vel": [h],
        "initialize_bonds": [h],
        "compute_bond_changes": [h, C.c_uint32],
        "prefix_sum_changes": [h, C.c_int],
        "counting_sort_changes": [h],
        "compute_particle_changes": [h, C.c_uint64],
        "physical_sort": [h],
        "inner_physical_loop": [h, C.c_float, C.c_float],
        "gene_action": [h],
        "step": [h, C.c_int, C.c_float, C.c_uint],
        "set_grid": [h, P(MgGrid)],
    }
    if prefix == "mg":
        sig["create"] = [P(h), C.c_int, C.c_int]
        sig["insert_particles"] = [h, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        sig["sync"] = [h]
        sig["fluid_setup"] = [h, C.c_int, C.c_int, P(C.c_int32), P(C.c_float), P(C.c_float),
                              P(C.c_float), P(C.c_float), C.c_int, C.c_int]
        sig["step_timed"] = [h, C.c_int, C.c_float, C.c_uint, P(C.c_double), P(C.c_int)]
        sig["get_params"] = [h, P(MgParams)]
        sig["get_grid"] = [h, P(MgGrid)]
        sig["debug_set_scan_ticket"] = [h, C.c_uint64]
        fl3 = P(C.c_float)
        sig["derive_params"] = [P(MgParams)] + [C.c_float] * 5 + [fl3, fl3] + [C.c_float] * 16
        sig["derive_grid"] = [P(MgGrid), fl3, fl3, C.c_float, C.c_float, C.c_float]
        sig["batch_create"] = [P(h), P(h), C.c_int]
        sig["batch_destroy"] = [h]
        sig["batch_step"] = [h, C.c_int, C.c_float, C.c_uint]
        sig["batch_sync"] = [h]
        sig["batch_classes"] = [h, P(C.c_int), P(C.c_int), P(C.c_uint)]
        sig["batch_autotune"] = [h, C.c_int, C.c_float, C.c_uint, P(C.c_uint)]
        sig["batch_step_timed"] = [h, C.c_int, C.c_float, C.c_uint, P(C.c_double), P(C.c_int)]
        sig["batch_launch_config"] = [h, P(C.c_int), P(C.c_uint)]
        sig["batch_class_config"] = [h, C.c_int, P(C.c_int), P(C.c_uint), P(C.c_int), P(C.c_int)]
        sig["batch_set_tile_shape"] = [h, C.c_int, C.c_int, C.c_int]
        sig["batch_set_persistent"] = [h, C.c_int, C.c_uint]
    else:
        sig["create"] = [P(h), C.c_int, C.c_int, C.c_int]
        sig["insert_particles"] = [h]
        sig["set_threads"] = [h, C.c_int]
    for name, args in sig.items():
        fn = getattr(lib, f"{prefix}_{name}")
        fn.argtypes = args
        fn.restype = C.c_int
    return lib
