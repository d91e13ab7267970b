"""The reference-signature set-up path on the GPU: the C++ facade's SetupGrid / FluidParamCL
and the C-ABI's mg_fluid_setup / mg_fluid_param store the MgGrid / MgParams that the oracle's
independent restatement derives (bit for bit), and a facade set up that way steps bit-equal to
the oracle. Also the cell scan's tile ticket across 2^32."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import fileio, scenes
from oracle import setup_ref
from tests.helpers import assert_bits_equal

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(REPO, "tests", "golden", "demo")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def facade_lib():
    lib = F.load_library()
    P, h, fl3 = C.POINTER, C.c_void_p, C.POINTER(C.c_float)
    lib.mgfs_create.argtypes = [P(h), C.c_int, C.c_int]
    lib.mgfs_destroy.argtypes = [h]
    lib.mgfs_set_host.argtypes = [h, C.c_int, C.c_void_p, C.c_size_t]
    lib.mgfs_get_host.argtypes = [h, C.c_int, C.c_void_p, C.c_size_t]
    lib.mgfs_setup_grid.argtypes = [h, fl3, fl3, C.c_float, C.c_float, C.c_float]
    lib.mgfs_fluid_param.argtypes = [h] + [C.c_float] * 5 + [fl3, fl3] + [C.c_float] * 16
    lib.mgfs_set_genome_dt.argtypes = [h, P(F.MgGenome), C.c_double]
    lib.mgfs_handle_of.argtypes = [h]
    lib.mgfs_handle_of.restype = h
    lib.mgfs_transfer_to.argtypes = [h]
    lib.mgfs_transfer_from.argtypes = [h]
    lib.mgfs_run.argtypes = [h, C.c_int, C.c_int]
    return lib


def raw(x):
    return bytes(memoryview(x))


def demo_case():
    sp = fileio.read_sim_params(os.path.join(DEMO, "SimParams.txt"))
    grid_args, param_args, _ = fileio.fluid_setup_args(sp)
    sc = fileio.load_scene(os.path.join(DEMO, "SimParams.txt"),
                           os.path.join(DEMO, "particles_pos_vel_color100001.csv"),
                           os.path.join(DEMO, "genome.csv"))
    return "demo", sc, grid_args, param_args, 40, True


def tissue_case():
    sc = scenes.tissue(n_side=10, b=16, seed=21)
    vmin, vmax = tuple(sc.grid.gmin), tuple(sc.grid.gmax)
    d = scenes.DEMO
    rest = scenes.rest_density(d["mass"], 0.5, d["ss"])
    grid_args = (vmin, vmax, 1.0, 1.0, 1.0)
    g = [d["grav"] * x for x in d["grav_dir"]]
    param_args = (d["ss"], d["sr"], d["pr"], d["mass"], rest, tuple(sc.params.pboundmin),
                  tuple(sc.params.pboundmax), d["estiff"], d["istiff"], d["visc"], d["surface_t"],
                  d["damp"], d["fmin"], d["fmax"], d["ffreq"], d["gslope"], g[0], g[1], g[2],
                  d["al"], d["vl"], 0.0, 1.0)
    return "tissue", sc, grid_args, param_args, 25, False


@pytest.mark.parametrize("case", [demo_case, tissue_case])
def test_facade_setup_path_matches_restatement_and_steps_like_oracle(built, case):
    from oracle.mg_oracle import Oracle

    name, sc, grid_args, param_args, steps, genes = case()
    say("case: facade set-up", name, sc.n, sc.b)
    lib = facade_lib()
    fl = C.c_float * 3
    want_g = setup_ref.grid_setup(*grid_args)
    want_p = setup_ref.params_from_fluid_param(*param_args)
    fh = C.c_void_p()
    assert lib.mgfs_create(C.byref(fh), sc.n, sc.b) == 0
    try:
        vmin, vmax, ss, cell, sr = grid_args
        assert lib.mgfs_setup_grid(fh, fl(*vmin), fl(*vmax), ss, cell, sr) == 0
        a = list(param_args)
        assert lib.mgfs_fluid_param(fh, *a[:5], fl(*a[5]), fl(*a[6]), *a[7:]) == 0
        assert lib.mgfs_set_genome_dt(fh, C.byref(sc.genome), sc.dt) == 0
        h = lib.mgfs_handle_of(fh)
        got_p, got_g = F.MgParams(), F.MgGrid()
        assert lib.mg_get_params(h, C.byref(got_p)) == 0
        assert lib.mg_get_grid(h, C.byref(got_g)) == 0
        assert raw(got_g) == raw(want_g), "facade SetupGrid"
        assert raw(got_p) == raw(want_p), "facade FluidParamCL"
        for b, arr in sc.buffers.items():
            x = np.ascontiguousarray(arr)
            assert lib.mgfs_set_host(fh, b, x.ctypes.data_as(C.c_void_p), x.nbytes) == 0
        assert lib.mgfs_transfer_to(fh) == 0
        assert lib.mgfs_run(fh, steps, int(genes)) == 0
        assert lib.mgfs_transfer_from(fh) == 0
        o = Oracle(sc.n, sc.b)
        o.set_grid(want_g)
        o.set_params(want_p)
        o.set_genome(sc.genome)
        o.TransferToCL(sc.buffers)
        o.Run2PhysicalSort()
        o.step(steps, sc.dt, F.STEP_GENES if genes else 0)
        for b in (F.FPOS, F.FVEL, F.FVEVAL, F.FPRESS, F.FDENSITY, F.FCONC, F.FEPIGEN, F.FELASTIDX,
                  F.FGCELL):
            shape, dt = F.buffer_shape(b, sc.n, sc.c, sc.b)
            x = np.empty(shape, dt)
            assert lib.mgfs_get_host(fh, b, x.ctypes.data_as(C.c_void_p), x.nbytes) == 0
            assert_bits_equal(x, o.download(b), f"{name} facade {F.BUFFER_NAMES[b]}")
    finally:
        lib.mgfs_destroy(fh)


def test_capi_fluid_setup_and_param_store_restatement(built):
    """mg_fluid_setup (FluidSetupCL's argument list) and mg_fluid_param (FluidParamCL's)."""
    from morphogenesis_amd.fluid_system import FluidSystem

    sp = fileio.read_sim_params(os.path.join(DEMO, "SimParams.txt"))
    grid_args, param_args, _ = fileio.fluid_setup_args(sp)
    want_g = setup_ref.grid_setup(*grid_args)
    want_p = setup_ref.params_from_fluid_param(*param_args)
    say("case: mg_fluid_setup / mg_fluid_param")
    s = FluidSystem(64, 6)
    lib = s.lib
    i3, f3 = C.c_int32 * 3, C.c_float * 3
    lib.mg_fluid_param.argtypes = [C.c_void_p] + [C.c_float] * 5 + [C.POINTER(C.c_float)] * 2 + [C.c_float] * 16
    assert lib.mg_fluid_setup(s.h, 64, want_g.srch, i3(*want_g.res), f3(*want_g.size),
                              f3(*want_g.delta), f3(*want_g.gmin), f3(*want_g.gmax),
                              want_g.total, want_g.chk) == 0
    a = list(param_args)
    assert lib.mg_fluid_param(s.h, *a[:5], f3(*a[5]), f3(*a[6]), *a[7:]) == 0
    got_p, got_g = F.MgParams(), F.MgGrid()
    assert lib.mg_get_params(s.h, C.byref(got_p)) == 0
    assert lib.mg_get_grid(s.h, C.byref(got_g)) == 0
    assert raw(got_g) == raw(want_g)
    assert raw(got_p) == raw(want_p)
    s.close()


def test_scan_ticket_across_32_bit_boundary(built):
    """The cell scan's tile ticket set just below 2^32 (a multiple of the tile count) and
    stepped across it: offsets and state stay bit-equal to the oracle."""
    from morphogenesis_amd.fluid_system import FluidSystem
    from oracle.mg_oracle import Oracle

    sc = scenes.tissue(n_side=14, b=8, seed=23, vmax=22.0)
    tiles = (sc.c + 4095) // 4096
    assert tiles >= 3 and (1 << 32) % tiles != 0
    seed_ticket = ((1 << 32) // tiles - 2) * tiles
    say("case: scan ticket", sc.c, "cells,", tiles, "tiles, ticket", seed_ticket)
    g = FluidSystem(sc.n, sc.b)
    o = Oracle(sc.n, sc.b)
    for sim in (g, o):
        sim.load_scene(sc)
        sim.Run2PhysicalSort()
    lib = g.lib
    assert lib.mg_debug_set_scan_ticket(g.h, seed_ticket + 1) != 0     # misaligned: refused
    assert lib.mg_debug_set_scan_ticket(g.h, seed_ticket) == 0
    for k in range(3):
        g.step(2, sc.dt, sc.flags)
        o.step(2, sc.dt, sc.flags)
        assert g.counts() == o.counts()
        for b in (F.FBIN_COUNT, F.FBIN_OFFSET, F.FGCELL, F.FPOS, F.FCONC):
            assert_bits_equal(g.download(b), o.download(b), f"ticket chunk {k}: {F.BUFFER_NAMES[b]}")
    g.close()
