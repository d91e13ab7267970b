"""Independent float32 restatement of the reference's host set-up arithmetic.

TEST INFRASTRUCTURE: the checker for the library's own derivation (mg_derive_params /
mg_derive_grid, used by mg_fluid_param, mg_fluid_setup and the C++ facade). The product never
imports this module.

  params_from_fluid_param  FluidParamCL, src/fluid_system.cpp:521-610
  grid_setup               SetupGrid, src/fluid_system.cpp:2073-2131, then FluidSetupCL's
                           gridScanMax / gridAdj, src/host_CL.cpp:39-106
Every operation is a float32 operation in the reference's order (numpy float32 scalars). The
reference's pow(float, float) resolves to the C library's powf, so pow here calls that same
library function (through ctypes, not through the product): its value is a property of the
host's libm, not of the reference's algorithm, and neither numpy's float32 power nor a
correctly rounded power reproduces it on every input (each can be an ulp away).
"""
from __future__ import annotations

import ctypes

import numpy as np

from morphogenesis_amd._ffi import MgGrid, MgParams


_libm = ctypes.CDLL("libm.so.6")
_libm.powf.restype = ctypes.c_float
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]


def _powf(x, y) -> np.float32:
    return np.float32(_libm.powf(float(np.float32(x)), float(np.float32(y))))


def params_from_fluid_param(ss, sr, pr, mass, rest, bmin, bmax, estiff, istiff, visc, surface_t,
                            damp, fmin, fmax, ffreq, gslope, gx, gy, gz, al, vl, a_f=0.0,
                            a_p=1.0) -> MgParams:
    """FluidParamCL (src/fluid_system.cpp:521-610) in float32."""
    F = np.float32
    p = MgParams()
    sr_, ss_ = F(sr), F(ss)
    p.psimscale = ss_
    p.psmoothradius = sr_
    p.pradius = F(pr)
    p.r2 = F(sr_ * sr_)
    p.pmass = F(mass)
    p.prest_dens = F(rest)
    for k in range(3):
        p.pboundmin[k] = F(bmin[k])
        p.pboundmax[k] = F(bmax[k])
    p.pextstiff, p.pintstiff, p.pvisc = F(estiff), F(istiff), F(visc)
    p.psurface_t, p.pdamp = F(surface_t), F(damp)
    p.pforce_min, p.pforce_max, p.pforce_freq, p.pground_slope = F(fmin), F(fmax), F(ffreq), F(gslope)
    p.pgravity[0], p.pgravity[1], p.pgravity[2] = F(gx), F(gy), F(gz)
    p.AL, p.VL = F(al), F(vl)
    p.AL2 = F(F(al) * F(al))
    p.VL2 = F(F(vl) * F(vl))
    pi = F(3.141592)
    p.pdist = _powf(F(p.pmass) / F(p.prest_dens), F(1) / F(3))
    p.poly6kern = F(F(315.0) / (F(64.0) * pi * _powf(sr_, F(9.0))))
    p.wendlandC2kern = F(F(21) / (F(2) * pi))
    p.spikykern = F(F(-45.0) / (pi * _powf(sr_, F(6.0))))
    p.lapkern = F(F(45.0) / (pi * _powf(sr_, F(6.0))))
    p.gausskern = F(F(1.0) / _powf(pi * F(2.0) * sr_ * sr_, F(1.5)))
    p.H = F(sr_ / ss_)
    p.d2 = F(ss_ * ss_)
    p.rd2 = F(F(p.r2) / F(p.d2))
    p.vterm = F(F(p.lapkern) * F(p.pvisc))
    p.actuation_factor, p.actuation_period = F(a_f), F(a_p)
    p.diff_scale = 1.0
    p.secrete_scale = 1.0
    p.freeze = 0
    p.frame = 0
    return p


def grid_setup(vmin, vmax, sim_scale, cell_size, smooth_radius) -> MgGrid:
    """SetupGrid (src/fluid_system.cpp:2073-2131) + FluidSetupCL (src/host_CL.cpp:39-106),
    float32 host arithmetic."""
    F = np.float32
    g = MgGrid()
    world_cellsize = F(F(cell_size) / F(sim_scale))
    res = []
    for k in range(3):
        size = F(F(vmax[k]) - F(vmin[k]))
        r = int(np.ceil(F(size / world_cellsize)))
        size = F(F(F(r) * F(cell_size)) / F(sim_scale))
        g.res[k] = r
        g.size[k] = size
        g.delta[k] = F(F(r) / size)
        g.gmin[k] = F(vmin[k])
        g.gmax[k] = F(vmax[k])
        res.append(r)
    gsrch = int(np.floor(F(F(F(2.0) * F(F(smooth_radius) / F(sim_scale))) / world_cellsize)) + F(1.0))
    gsrch = max(gsrch, 2)
    for k in range(3):
        g.scan_max[k] = res[k] - gsrch
    g.total = res[0] * res[1] * res[2]
    g.srch = gsrch
    g.adj_cnt = gsrch ** 3
    g.chk = 0
    cell = 0
    for y in range(gsrch):
        for z in range(gsrch):
            for x in range(gsrch):
                g.adj[cell] = (y * res[2] + z) * res[0] + x
                cell += 1
    return g
