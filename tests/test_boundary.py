"""The C-ABI boundary on the CPU: the library loads, exports every entry point the headers
declare, and the Python mirror keeps the reference's method names. No compute calls."""
import ctypes as C
import os
import re

import pytest

from morphogenesis_amd import _ffi as F

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(mg_[a-z_0-9]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(F.LIB_PATH):
        pytest.skip("libmg_sph.so not built")
    lib = C.CDLL(F.LIB_PATH)
    names = declared("mg_sph.h")
    assert "mg_step" in names and "mg_upload" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    for n in ("mgfs_create", "mgfs_run", "mgfs_transfer_to", "mgfs_transfer_from"):
        assert hasattr(lib, n), n


def test_struct_layouts_match_header():
    # sizes fixed by include/mg_types.h (FGenome layout of src/fluid.h:398-422)
    assert C.sizeof(F.MgGenome) == 4 * (16 + 16 + 256 + 16 + 16 + 16 * 33 * 2 + 24 + 36)
    assert C.sizeof(F.MgGrid) == 4 * (3 + 3 + 12 + 4 + 64)
    assert C.sizeof(F.MgParams) == 4 * (15 + 9 + 18 + 2)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(F.MgError):
        F.load_library(str(tmp_path / "nope.so"))


def test_python_mirror_has_reference_method_names():
    from morphogenesis_amd.fluid_system import FluidSystem

    for m in ("Run2PhysicalSort", "Run2InnerPhysicalLoop", "Run2GeneAction", "InsertParticlesCL",
              "PrefixSumCellsCL", "CountingSortFullCL", "ComputePressureCL", "ComputeForceCL",
              "AdvanceCL", "ComputeDiffusionCL", "ComputeGenesCL", "ZeroVelCL", "TransferToCL",
              "TransferFromCL", "setFreeze"):
        assert callable(getattr(FluidSystem, m)), m


def test_buffer_sizes_match_reference_layout():
    n, c, b = 100, 27, 6
    # src/fluid.h strides: float4 vectors, uint[B][9] bond records, uint[B][2] incoming,
    # float[16] conc (AoS), uint[16] epigen (SoA)
    expect = {F.FPOS: n * 16, F.FELASTIDX: n * b * 9 * 4, F.FPARTICLEIDX: n * b * 2 * 4,
              F.FCONC: n * 64, F.FEPIGEN: n * 64, F.FBIN_COUNT: c * 4,
              F.FBIN_COUNT_ACTIVE_GENES: 16 * c * 4}
    for buf, nbytes in expect.items():
        shape, dt = F.buffer_shape(buf, n, c, b)
        assert int(__import__("numpy").prod(shape)) * 4 == nbytes


def _struct_bytes(x):
    return bytes(memoryview(x))


def _param_cases():
    import numpy as np

    from morphogenesis_amd import scenes
    from morphogenesis_amd.fileio import fluid_setup, read_sim_params

    cases = []
    d = dict(scenes.DEMO)
    for sr in (1.0, 0.75, 1.5, 2.0):
        for spacing in (0.5, 0.37):
            rest = scenes.rest_density(d["mass"], spacing, d["ss"])
            cases.append((d["ss"], sr, d["pr"], d["mass"], rest, (2.0, 2.0, 2.0), (34.0, 30.0, 34.0),
                          d["estiff"], d["istiff"], d["visc"], d["surface_t"], d["damp"], d["fmin"],
                          d["fmax"], d["ffreq"], d["gslope"], 0.0, -2.0, 0.0, d["al"], d["vl"], 0.0, 1.0))
    rng = np.random.default_rng(7)
    for _ in range(200):
        v = [float(np.float32(x)) for x in rng.uniform(0.05, 9.0, 27)]
        v[0] = float(np.float32(rng.choice([1.0, 0.5, 0.25, 1.3, 0.004])))
        cases.append(tuple(v[:5]) + (tuple(v[5:8]), tuple(v[8:11])) + tuple(v[11:27]))
    sp = read_sim_params(os.path.join(REPO, "tests", "golden", "demo", "SimParams.txt"))
    return cases, fluid_setup, sp


def test_setup_derivation_matches_independent_restatement(built):
    """mg_derive_params / mg_derive_grid (the code behind mg_fluid_param, mg_fluid_setup and the
    C++ facade's FluidParamCL / SetupGrid) are bit-equal to the oracle's independent float32
    restatement of FluidParamCL (src/fluid_system.cpp:521-610) and SetupGrid + FluidSetupCL
    (src/fluid_system.cpp:2073-2131, src/host_CL.cpp:39-106)."""
    import numpy as np

    from oracle import setup_ref

    cases, fluid_setup, sp = _param_cases()
    for args in cases:
        a = list(args)
        flat = a[:5] + [a[5], a[6]] + a[7:]
        want = setup_ref.params_from_fluid_param(*flat)
        got = F.params_from_fluid_param(*flat)
        assert _struct_bytes(got) == _struct_bytes(want), args
    rng = np.random.default_rng(11)
    grids = [((0, 0, 0), (36, 36, 36), 1.0, 1.0, 1.0), ((0, 0, 0), (20, 20, 20), 1.0, 1.0, 1.0),
             ((-3.5, 0, 1.25), (17.0, 9.5, 30.0), 1.0, 0.75, 0.75),
             ((0, 0, 0), (21, 21, 21), 1.0, 1.5, 1.5), ((0, 0, 0), (12, 12, 12), 0.5, 1.0, 1.0)]
    for _ in range(200):
        lo = [float(np.float32(x)) for x in rng.uniform(-10, 10, 3)]
        hi = [float(np.float32(l + x)) for l, x in zip(lo, rng.uniform(4, 60, 3))]
        sr = float(np.float32(rng.uniform(0.3, 2.5)))
        ss = float(np.float32(rng.choice([1.0, 0.5, 2.0, 0.9])))
        cell = float(np.float32(2 * sr / rng.choice([2.0, 1.5, 3.0])))
        grids.append((lo, hi, ss, cell, sr))
    for (lo, hi, ss, cell, sr) in grids:
        want = setup_ref.grid_setup(lo, hi, ss, cell, sr)
        if want.srch ** 3 > 64:
            continue
        got = F.grid_setup(lo, hi, ss, cell, sr)
        assert _struct_bytes(got) == _struct_bytes(want), (lo, hi, ss, cell, sr)
    # the reference's demo SimParams through the file set-up path
    params, grid, _ = fluid_setup(sp)
    P = sp.param
    assert grid.srch == 3 and list(grid.res) == [int(r) for r in grid.res]
    assert params.psmoothradius == np.float32(P["PSMOOTHRADIUS"])


def test_derive_entry_points_reject_bad_arguments(built):
    lib = F.load_library()
    g = F.MgGrid()
    fl = C.c_float * 3
    assert lib.mg_derive_grid(C.byref(g), fl(0, 0, 0), fl(10, 10, 10), 1.0, 0.0, 1.0) == -1
    assert lib.mg_derive_grid(C.byref(g), fl(0, 0, 0), fl(10, 10, 10), 1.0, 0.2, 1.0) == -1  # 21^3 > 64
    assert lib.mg_derive_grid(None, fl(0, 0, 0), fl(10, 10, 10), 1.0, 1.0, 1.0) == -1
