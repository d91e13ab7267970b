"""Shared test helpers: hand-built scenes and comparisons."""
from __future__ import annotations

import numpy as np

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import scenes


def mini_scene(pos, b=4, vmax=12.0, bonds=(), genome=None, conc=None, epigen=None, vel=None,
               params_over=None, spacing=0.5, sr=1.0, vmin=0.0):
    """A scene from explicit positions. bonds: (owner, partner, rest, modulus, damping, limit).
    sr is the smoothing radius; the grid cell is 2*sr/PGRID_DENSITY = sr, as the reference's
    set-up derives it, so the 27-cell search covers the density cut-off r < sr."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    vmin3, vmax3 = (vmin, vmin, vmin), (vmax, vmax, vmax)
    over = dict(params_over or {})
    over.setdefault("sr", sr)
    params = scenes.make_params(vmin3, vmax3, spacing=spacing, **over)
    grid = scenes.make_grid(vmin3, vmax3, sr=over["sr"])
    bufs = scenes.base_buffers(n, b, pos)
    rec = bufs[F.FELASTIDX]
    pidx = bufs[F.FPARTICLEIDX]
    used = np.zeros(n, int)
    inc = np.zeros(n, int)
    f = lambda x: np.float32(x).view(np.uint32)
    for (i, j, rest, k, c, lim) in bonds:
        a = used[i]
        used[i] += 1
        rec[i, a] = [j, f(lim), f(rest), f(k), f(c), bufs[F.FPARTICLE_ID][j], f(0), f(0), 0]
        pidx[j, inc[j]] = [i, a]
        inc[j] += 1
    if conc is not None:
        bufs[F.FCONC][:] = np.asarray(conc, np.float32).reshape(n, F.NUM_TF)
    if epigen is not None:
        bufs[F.FEPIGEN][:] = np.asarray(epigen, np.uint32).reshape(F.NUM_GENES, n)
    if vel is not None:
        v = np.asarray(vel, np.float32).reshape(n, 3)
        bufs[F.FVEL][:, :3] = v
        bufs[F.FVEVAL][:, :3] = v
    return scenes.Scene("mini", n, b, params, grid, genome or scenes.empty_genome(), bufs,
                        1e-4, 1, 0, {})


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def assert_bits_equal(a, b, what=""):
    ba, bb = bits(a), bits(b)
    if not np.array_equal(ba, bb):
        diff = np.argwhere(ba != bb)
        first = tuple(diff[0]) if diff.size else ()
        raise AssertionError(f"{what}: {len(diff)} words differ; first at {first}: "
                             f"{a[first] if first else ''} vs {b[first] if first else ''}")


def per_cell_sets(gcell, pid, num_points):
    cells = {}
    for c, p in zip(gcell[:num_points].tolist(), pid[:num_points].tolist()):
        cells.setdefault(c, set()).add(p)
    return cells
