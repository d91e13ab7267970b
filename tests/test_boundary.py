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
