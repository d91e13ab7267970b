"""mg_prefix_sum_cells_noalloc (the cell scan with its scratch in a fixed device array, for
callers that cannot allocate) against mg_prefix_sum_cells on the same counts: bit-exact, on
two calls in a row, the second with a smaller input; and against numpy on a 5M-cell array."""
import ctypes as C
import sys

import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hip_runtime():
    """The HIP runtime libmg_sph.so links (the same library instance): the test's own device
    buffers come from it; the entry point under test allocates nothing."""
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)  # already loaded by libmg_sph.so
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


def noalloc(lib, hip, cnt_host):
    cnt_host = np.ascontiguousarray(cnt_host, np.uint32)
    n = cnt_host.size
    bufs = [C.c_void_p() for _ in range(3)]
    for b, size in zip(bufs, (4 * n, 4 * n, 8)):
        assert hip.hipMalloc(C.byref(b), size) == 0
    try:
        assert hip.hipMemcpy(bufs[0], cnt_host.ctypes.data_as(C.c_void_p), 4 * n, H2D) == 0
        junk = np.full(n, 0xDEADBEEF, np.uint32)
        assert hip.hipMemcpy(bufs[1], junk.ctypes.data_as(C.c_void_p), 4 * n, H2D) == 0
        rc = lib.mg_prefix_sum_cells_noalloc(bufs[0], bufs[1], n, bufs[2], None)
        assert hip.hipDeviceSynchronize() == 0
        off = np.empty(n, np.uint32)
        rep = np.empty(2, np.uint32)
        assert hip.hipMemcpy(off.ctypes.data_as(C.c_void_p), bufs[1], 4 * n, D2H) == 0
        assert hip.hipMemcpy(rep.ctypes.data_as(C.c_void_p), bufs[2], 8, D2H) == 0
        return rc, off, rep
    finally:
        for b in bufs:
            hip.hipFree(b)


def test_noalloc_scan_matches_prefix_sum_cells(built):
    from morphogenesis_amd.fluid_system import FluidSystem

    lib = F.load_library()
    hip = hip_runtime()
    lib.mg_prefix_sum_cells_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    lib.mg_prefix_sum_cells_noalloc.restype = C.c_int
    for n_side, seed in ((40, 61), (12, 62)):      # 13,824 cells (4 tiles), then 1,000 (1 tile)
        sc = scenes.tissue(n_side=n_side, b=8, seed=seed)
        say("case: noalloc scan vs mg_prefix_sum_cells,", sc.c, "cells")
        g = FluidSystem(sc.n, sc.b)
        g.load_scene(sc)
        _, _, gcnt = g.InsertParticlesCL(readback=True)
        g.PrefixSumCellsCL(1)
        want = g.download(F.FBIN_OFFSET)
        rc, got, rep = noalloc(lib, hip, gcnt)
        assert rc == 0, lib.mg_last_error()
        np.testing.assert_array_equal(got, want)
        assert rep[0] == gcnt.sum() and rep[1] == 0
        g.close()
    # many tiles, against numpy
    rng = np.random.default_rng(7)
    cnt = rng.integers(0, 40, 5_000_000).astype(np.uint32)
    say("case: noalloc scan, 5,000,000 cells")
    rc, got, rep = noalloc(lib, hip, cnt)
    assert rc == 0
    np.testing.assert_array_equal(got, np.concatenate([[0], np.cumsum(cnt, dtype=np.uint64)[:-1]]).astype(np.uint32))
    assert rep[0] == cnt.sum() and rep[1] == 0
    # beyond the fixed scratch: refused before anything runs
    r = lib.mg_prefix_sum_cells_noalloc(C.c_void_p(16), C.c_void_p(16), 16384 * 4096 + 1, None, None)
    assert r == -5 and b"exceeds" in lib.mg_last_error()
