"""Synthetic scenes in the reference buffer layout.

Each scene is the state a caller of the reference's FluidSystem would upload with
TransferToCL: parameters (FluidParamCL), grid (SetupGrid + FluidSetupCL), genome
(FGenome) and the per-particle buffers of src/fluid.h:138-206. Units follow the
reference's demo (src/demo/SimParams.txt): PSIMSCALE = 1, PSMOOTHRADIUS = 1, grid cell =
2*sr/PGRID_DENSITY = 1, so "h" in a workload description is one world unit.

Workloads (seeded):
  fluid_block   32^3 particles, spacing 0.5, jitter +-0.05, VOL [0,20]^3, no bonds/genes.
  tissue        n^3 particles (64^3 = 262,144 for the headline workload), spacing 0.5,
                jitter +-0.05, bonds to every neighbour within 1.5 x spacing (nearest first,
                one directed bond per pair, up to B slots per record), rest length = initial
                distance, modulus ~ U[0.5, 2] * k0, damping ratio 0.1, 16 gene flags
                (FEPIGEN = INT_MAX with p = 0.3) and 16 concentrations ~ U[0, 1].
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import _ffi as F

INT_MAX = 2147483647


@dataclasses.dataclass
class Scene:
    name: str
    n: int
    b: int
    params: F.MgParams
    grid: F.MgGrid
    genome: F.MgGenome
    buffers: dict
    dt: float
    steps: int
    flags: int
    info: dict = dataclasses.field(default_factory=dict)

    @property
    def c(self) -> int:
        return int(self.grid.total)


# Reference demo parameters (src/demo/SimParams.txt), PSMOOTHRADIUS = PSIMSCALE = 1.
DEMO = dict(ss=1.0, sr=1.0, pr=1.0, mass=0.00020543, visc=0.1, surface_t=0.1, istiff=1.0,
            estiff=50000.0, damp=100.0, al=150.0, vl=3.0, grav=2.0, grav_dir=(0.0, -1.0, 0.0),
            fmin=0.0, fmax=0.0, ffreq=16.0, gslope=0.0, grid_density=2.0)


def rest_density(mass: float, spacing: float, ss: float = 1.0) -> float:
    """SetupSpacing (src/fluid_system.cpp:2531-2535): density from spacing."""
    f = np.float32
    pdist = f(f(spacing) * f(ss) / f(0.87))
    return float(f(f(mass) / F.powf(pdist, 3.0)))


def make_params(vmin, vmax, spacing=0.5, **over) -> F.MgParams:
    d = dict(DEMO)
    d.update(over)
    f = np.float32
    grid_size = float(f(f(2.0) * f(d["sr"]) / f(d["grid_density"])))
    # PBOUNDMIN/MAX = VOL -/+ 2*gridsize/simscale (src/fluid_system.cpp:2539-2542)
    off = 2.0 * (grid_size / d["ss"])
    bmin = [float(f(v + off)) for v in vmin]
    bmax = [float(f(v - off)) for v in vmax]
    rest = d.get("rest", rest_density(d["mass"], spacing, d["ss"]))
    g = [d["grav"] * x for x in d["grav_dir"]]
    p = F.params_from_fluid_param(d["ss"], d["sr"], d["pr"], d["mass"], rest, bmin, bmax,
                                  d["estiff"], d["istiff"], d["visc"], d["surface_t"], d["damp"],
                                  d["fmin"], d["fmax"], d["ffreq"], d["gslope"], g[0], g[1], g[2],
                                  d["al"], d["vl"])
    return p


def make_grid(vmin, vmax, ss=1.0, sr=1.0, grid_density=2.0) -> F.MgGrid:
    f = np.float32
    cell = float(f(f(2.0) * f(sr) / f(grid_density)))
    return F.grid_setup(vmin, vmax, ss, cell, sr)


def empty_genome() -> F.MgGenome:
    g = F.MgGenome()
    for i in range(F.NUM_GENES):
        g.delay[i] = 1
    for t in range(F.NUM_TF):
        g.tf_breakdown_rate[t] = 1
    return g


def example_genome() -> F.MgGenome:
    """SetupExampleGenome (src/fluid_system.cpp:2415-2519)."""
    g = F.MgGenome()
    for i in range(F.NUM_GENES):
        g.delay[i] = 1
        for j in range(F.NUM_GENES):
            g.sensitivity[i][j] = j
    for t in range(F.NUM_TF):
        g.tf_diffusability[t] = 0 if t < F.NUM_TF // 2 else 1
        g.tf_breakdown_rate[t] = 1
    g.secrete[0][2 * F.NUM_TF] = 2
    g.secrete[0][0], g.secrete[0][1] = 1, 1
    g.secrete[0][2], g.secrete[0][3] = 3, 4
    g.activate[0][2 * F.NUM_GENES] = 1
    g.activate[0][0], g.activate[0][1] = 5, 6
    # remodelling rows elastin, collagen, apatite (src/fluid_system.cpp:2441-2486), columns in
    # FGenome::params order: elongation_threshold, elongation_factor, strength_threshold,
    # strengthening_factor, max/min_rest_length, max/min_modulus, elastLim,
    # default_rest_length, default_modulus, default_damping
    rows = [(0.5, 0.002, 1.0, 0.002, 1.0, 0.3, 1e5, 10, 8, 0.5, 1000, 10),
            (4.0, 0.01, 4.1, 0.01, 1.0, 0.3, 1e9, 1e6, 0.55, 0.5, 1e7, 100),
            (1.0, 0.001, 1.0, 0.001, 1.0, 0.3, 1e9, 1e6, 0.05, 0.5, 1e7, 1000)]
    for r, vals in enumerate(rows):
        for k, v in enumerate(vals):
            g.param[r][k] = v
    # tanh remodelling curve of elastin's strengthening (src/fluid_system.cpp:2489-2497);
    # collagen and apatite rows stay zero
    for k, v in enumerate((0.0, 0.0, 0.0, 0.0, 1.008, 0.01, 10.0, 7.0)):
        g.tanh_param[0][k] = v
    return g


def tissue_genome(seed: int) -> F.MgGenome:
    """A seeded genome for the tissue workload: sparse 0/1 sensitivities, every gene
    secretes one TF and activates the next gene above a threshold; TFs 8..15 diffuse."""
    rng = np.random.default_rng(seed + 7)
    g = F.MgGenome()
    for i in range(F.NUM_GENES):
        g.delay[i] = int(rng.integers(1, 4))
        for j in range(F.NUM_GENES):
            g.sensitivity[i][j] = int(rng.random() < 0.25)
        g.secrete[i][0] = int(rng.integers(0, F.NUM_TF))
        g.secrete[i][1] = int(rng.integers(-1, 3))
        g.secrete[i][2 * F.NUM_TF] = 1
        g.activate[i][0] = (i + 1) % F.NUM_GENES
        g.activate[i][1] = 2
        g.activate[i][2 * F.NUM_GENES] = 1
    for t in range(F.NUM_TF):
        g.tf_diffusability[t] = 0 if t < F.NUM_TF // 2 else 1
        g.tf_breakdown_rate[t] = 1
    return g


def base_buffers(n: int, b: int, pos: np.ndarray) -> dict:
    """Per-particle buffers as SetupAddVolumeMorphogenesis2 / AddParticleMorphogenesis2 fill
    them (src/fluid_system.cpp:1037-1167): zero velocity, IDs from 1, empty bond slots."""
    u32, f32 = np.uint32, np.float32
    bufs = {}
    p4 = np.zeros((n, 4), f32)
    p4[:, :3] = pos
    bufs[F.FPOS] = p4
    bufs[F.FVEL] = np.zeros((n, 4), f32)
    bufs[F.FVEVAL] = np.zeros((n, 4), f32)
    bufs[F.FFORCE] = np.zeros((n, 4), f32)
    bufs[F.FPRESS] = np.zeros(n, f32)
    bufs[F.FDENSITY] = np.zeros(n, f32)
    bufs[F.FAGE] = np.zeros(n, u32)
    bufs[F.FCOLOR] = (np.arange(n, dtype=np.uint64) * 2654435761 % (1 << 32)).astype(u32)
    bufs[F.FGCELL] = np.full(n, F.GRID_UNDEF, u32)
    bufs[F.FGNDX] = np.zeros(n, u32)
    pid = np.arange(1, n + 1, dtype=u32)
    bufs[F.FPARTICLE_ID] = pid
    bufs[F.FNERVEIDX] = np.where(pid % 10 == 0, pid // 10, 0).astype(u32)
    mass_radius = (int(0.00020543 * 255.0 * 255.0) << 16) | int(1.0 * 255.0 * 255.0)
    bufs[F.FMASS_RADIUS] = np.full(n, mass_radius, u32)
    bufs[F.FBIN] = np.zeros(n, u32)
    rec = np.zeros((n, b, F.DATA_PER_BOND), u32)
    rec[:, :, 0] = F.NONE
    rec[:, :, 5] = F.NONE
    bufs[F.FELASTIDX] = rec
    bufs[F.FPARTICLEIDX] = np.full((n, b, 2), F.NONE, u32)
    bufs[F.FCONC] = np.zeros((n, F.NUM_TF), f32)
    bufs[F.FEPIGEN] = np.zeros((F.NUM_GENES, n), u32)
    return bufs


def lattice(nx: int, ny: int, nz: int, spacing: float, origin, jitter: float, rng) -> np.ndarray:
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pos = np.stack([ix, iy, iz], axis=-1).reshape(-1, 3).astype(np.float64) * spacing
    pos += np.asarray(origin, np.float64)
    pos += rng.uniform(-jitter, jitter, size=pos.shape)
    return pos.astype(np.float32)


def add_bonds(bufs: dict, pos: np.ndarray, b: int, radius: float, k0: float, zeta: float,
              limit: float, mass: float, rng) -> dict:
    """One directed bond per neighbour pair within `radius`, owned by the lower index, listed
    nearest first; FPARTICLEIDX of the partner gets the reciprocal (owner, slot) entry
    (src/fluid_system_cuda.cu:1892-1895, 1914-1917 describe the reciprocity). Damping is
    zeta * 2*sqrt(m*k), the reference's "optimal damping" convention
    (src/fluid_system_cuda.cu:3280-3281)."""
    from scipy.spatial import cKDTree

    n = pos.shape[0]
    rec = bufs[F.FELASTIDX]
    pidx = bufs[F.FPARTICLEIDX]
    if b == 0:
        return {"bonds": 0}
    tree = cKDTree(pos.astype(np.float64))
    pairs = tree.query_pairs(r=radius, output_type="ndarray")
    if pairs.size == 0:
        return {"bonds": 0}
    i, j = pairs[:, 0], pairs[:, 1]
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    d = pos[lo].astype(np.float32) - pos[hi].astype(np.float32)
    dist = np.sqrt((d.astype(np.float32) ** 2).sum(axis=1, dtype=np.float32)).astype(np.float32)
    order = np.lexsort((hi, dist, lo))  # by owner, then distance, then partner
    lo, hi, dist = lo[order], hi[order], dist[order]
    # slot index within owner
    start = np.searchsorted(lo, lo, side="left")
    slot = np.arange(lo.size) - start
    keep = slot < b
    lo, hi, dist, slot = lo[keep], hi[keep], dist[keep], slot[keep]
    modulus = (rng.uniform(0.5, 2.0, size=lo.size) * k0).astype(np.float32)
    damping = (zeta * 2.0 * np.sqrt(mass * modulus.astype(np.float64))).astype(np.float32)
    f = lambda a: np.asarray(a, np.float32).view(np.uint32)
    rec[lo, slot, 0] = hi.astype(np.uint32)
    rec[lo, slot, 1] = f(np.full(lo.size, limit))
    rec[lo, slot, 2] = f(dist)
    rec[lo, slot, 3] = f(modulus)
    rec[lo, slot, 4] = f(damping)
    rec[lo, slot, 5] = bufs[F.FPARTICLE_ID][hi]
    rec[lo, slot, 6] = f(np.zeros(lo.size))
    rec[lo, slot, 7] = f(np.zeros(lo.size))
    rec[lo, slot, 8] = 0
    # incoming lists, ascending (owner, slot)
    o2 = np.lexsort((slot, lo, hi))
    hi2, lo2, slot2 = hi[o2], lo[o2], slot[o2]
    st = np.searchsorted(hi2, hi2, side="left")
    a = np.arange(hi2.size) - st
    ok = a < b
    pidx[hi2[ok], a[ok], 0] = lo2[ok].astype(np.uint32)
    pidx[hi2[ok], a[ok], 1] = slot2[ok].astype(np.uint32)
    per = np.bincount(lo, minlength=n)
    inc = np.bincount(hi, minlength=n)
    return {"bonds": int(lo.size), "max_out": int(per.max()), "max_in": int(inc.max()),
            "mean_bonds_per_particle": float((per + inc).mean())}


def add_genes(bufs: dict, n: int, p_active: float, rng, conc: bool = True):
    epi = bufs[F.FEPIGEN]
    act = rng.random((F.NUM_GENES, n)) < p_active
    epi[act] = INT_MAX
    if conc:
        bufs[F.FCONC][:] = rng.uniform(0.0, 1.0, size=(n, F.NUM_TF)).astype(np.float32)


def fluid_block(n_side: int = 32, seed: int = 1, spacing: float = 0.5, jitter: float = 0.05,
                vmax: float = 20.0, steps: int = 100, dt: float = 1e-4) -> Scene:
    """Plain SPH fluid block: n_side^3 particles, no bonds, no genes, gravity only."""
    rng = np.random.default_rng(seed)
    vmin3, vmax3 = (0.0, 0.0, 0.0), (vmax, vmax, vmax)
    params = make_params(vmin3, vmax3, spacing=spacing)
    grid = make_grid(vmin3, vmax3)
    nx = ny = nz = n_side
    extent = [(m - 1) * spacing for m in (nx, ny, nz)]
    lo = [(p + q) / 2 - e / 2 for p, q, e in zip(params.pboundmin, params.pboundmax, extent)]
    lo[1] = params.pboundmin[1] + 0.25
    pos = lattice(nx, ny, nz, spacing, lo, jitter, rng)
    n = pos.shape[0]
    bufs = base_buffers(n, 0, pos)
    return Scene(f"fluid_block_{n_side}", n, 0, params, grid, empty_genome(), bufs, dt, steps, 0,
                 {"spacing": spacing, "jitter": jitter, "seed": seed})


def tissue(n_side: int = 64, b: int = 32, seed: int = 1, spacing: float = 0.5,
           jitter: float = 0.05, vmax: float | None = None, k0: float = 10.0, zeta: float = 0.1,
           limit: float = 2.0, genes: bool = True, steps: int = 1000, dt: float = 1e-4,
           diff_scale: float = 0.01, secrete_scale: float = 0.001, reserve: int = 0,
           dims: tuple | None = None) -> Scene:
    """Bonded elastic tissue block (the headline workload at n_side = 64, B = 32), or an
    nx x ny x nz block when `dims` is given. `reserve` extra null particles (AddNullPoints,
    src/fluid_system.cpp:988-1034) are parked at PBOUNDMAX, outside the scan window."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims if dims is not None else (n_side, n_side, n_side)
    if vmax is None:
        vmax = float(np.ceil((max(nx, ny, nz) - 1) * spacing + 4.5))
    vmin3, vmax3 = (0.0, 0.0, 0.0), (vmax, vmax, vmax)
    params = make_params(vmin3, vmax3, spacing=spacing)
    params.diff_scale = diff_scale
    params.secrete_scale = secrete_scale
    grid = make_grid(vmin3, vmax3)
    extent = [(m - 1) * spacing for m in (nx, ny, nz)]
    lo = [(p + q) / 2 - e / 2 for p, q, e in zip(params.pboundmin, params.pboundmax, extent)]
    lo[1] = params.pboundmin[1] + 0.25
    pos = lattice(nx, ny, nz, spacing, lo, jitter, rng)
    n_live = pos.shape[0]
    if reserve:
        park = np.tile(np.asarray(params.pboundmax, np.float32), (reserve, 1))
        pos = np.concatenate([pos, park])
    n = pos.shape[0]
    bufs = base_buffers(n, b, pos)
    info = add_bonds(bufs, pos[:n_live], b, 1.5 * spacing, k0, zeta, limit, params.pmass, rng)
    genome = empty_genome()
    flags = 0
    if genes:
        add_genes(bufs, n, 0.3, rng)
        if reserve:
            bufs[F.FEPIGEN][:, n_live:] = 0
            bufs[F.FCONC][n_live:] = 0
        genome = tissue_genome(seed)
        flags = F.STEP_GENES
    info.update({"spacing": spacing, "jitter": jitter, "seed": seed, "k0": k0, "zeta": zeta,
                 "limit": limit, "reserve": reserve})
    name = f"tissue_{n_side}_B{b}" if dims is None else f"tissue_{nx}x{ny}x{nz}_B{b}"
    return Scene(name, n, b, params, grid, genome, bufs, dt, steps, flags, info)


def mixed(n_total: int = 4194304, b: int = 32, seed: int = 2, slab_frac: float = 0.7,
          spacing: float = 0.5, droplet_frac: float = 0.05, k0: float = 10.0, zeta: float = 0.1,
          limit: float = 2.0, steps: int = 200, dt: float = 1e-4) -> Scene:
    """A bonded slab (slab_frac of the particles, a square nx x ny x nx jittered lattice
    resting on the floor, B bond slots) with a fluid column dropped onto it: a cylinder above
    the slab centre at the slab's spacing falling at 1 unit/s, with droplet_frac of the fluid
    packed into dense droplets at half the spacing (about 64 particles per cell), so cell
    occupancy runs from 1 to about 64. Genes and concentrations on every particle. Particle
    counts are exact: the fluid takes n_total - n_slab."""
    rng = np.random.default_rng(seed)
    n_slab_target = int(round(slab_frac * n_total))
    nx = int(np.sqrt(n_slab_target / 45.0))
    ny = max(2, int(round(n_slab_target / (nx * nx))))
    n_slab = nx * nx * ny
    n_fluid = n_total - n_slab
    n_drop = int(round(droplet_frac * n_fluid))
    n_col = n_fluid - n_drop
    side = (nx - 1) * spacing
    radius = 0.35 * side
    col_h = n_col / ((np.pi * radius * radius) / spacing ** 3)
    vxz = float(np.ceil(side + 5.5))
    vy = float(np.ceil(2.5 + (ny - 1) * spacing + 2.0 + col_h + 6.0))
    vmin3, vmax3 = (0.0, 0.0, 0.0), (vxz, vy, vxz)
    params = make_params(vmin3, vmax3, spacing=spacing)
    params.diff_scale = 0.01
    params.secrete_scale = 0.001
    grid = make_grid(vmin3, vmax3)
    cx = (params.pboundmin[0] + params.pboundmax[0]) / 2
    cz = (params.pboundmin[2] + params.pboundmax[2]) / 2
    y0 = params.pboundmin[1] + 0.25
    slab = lattice(nx, ny, nx, spacing, (cx - side / 2, y0, cz - side / 2), 0.05, rng)
    top = y0 + (ny - 1) * spacing + 2.0
    # column: jittered lattice points inside the cylinder, lowest first, exactly n_col of them
    m = int(np.ceil(2 * radius / spacing)) + 1
    k = int(np.ceil(col_h / spacing)) + 3
    g = lattice(m, k, m, spacing, (cx - radius, top, cz - radius), 0.05, rng)
    inside = (g[:, 0] - cx) ** 2 + (g[:, 2] - cz) ** 2 <= radius * radius
    g = g[inside]
    g = g[np.argsort(g[:, 1], kind="stable")][:n_col]
    # droplets: dense balls at half the spacing near the top of the column
    drops = []
    left = n_drop
    while left > 0:
        r = 2.0
        c = np.array([cx + rng.uniform(-0.5, 0.5) * radius, top + col_h + rng.uniform(1.0, 3.0),
                      cz + rng.uniform(-0.5, 0.5) * radius])
        q = int(np.ceil(2 * r / (spacing / 2))) + 1
        ball = lattice(q, q, q, spacing / 2, c - r, 0.01, rng)
        ball = ball[((ball - c) ** 2).sum(1) <= r * r][:left]
        drops.append(ball)
        left -= ball.shape[0]
    fluid = np.concatenate([g] + drops).astype(np.float32)
    pos = np.concatenate([slab, fluid]).astype(np.float32)
    assert pos.shape[0] == n_total, (pos.shape, n_total)
    bufs = base_buffers(n_total, b, pos)
    bufs[F.FVEL][n_slab:, 1] = -1.0
    bufs[F.FVEVAL][n_slab:, 1] = -1.0
    info = add_bonds(bufs, slab, b, 1.5 * spacing, k0, zeta, limit, params.pmass, rng)
    add_genes(bufs, n_total, 0.3, rng)
    info.update({"n_slab": n_slab, "n_fluid": n_fluid, "n_droplets": n_drop, "slab": (nx, ny, nx),
                 "column_radius": radius, "column_height": col_h, "seed": seed})
    return Scene(f"mixed_{n_total}", n_total, b, params, grid, tissue_genome(seed), bufs, dt, steps,
                 F.STEP_GENES, info)


def batch_sizes(count: int = 256, lo: int = 2048, hi: int = 65536, seed: int = 3) -> list:
    """Particle counts drawn log-uniformly in [lo, hi] (seeded), each rounded to a block
    nx x ny x nz of the tissue lattice (nx = ny, nz chosen to land closest)."""
    rng = np.random.default_rng(seed)
    out = []
    for v in np.exp(rng.uniform(np.log(lo), np.log(hi), count)):
        nx = max(2, int(round(v ** (1.0 / 3.0))))
        nz = max(2, int(round(v / (nx * nx))))
        out.append((nx, nx, nz))
    return out


def tissue_batch(count: int = 256, b: int = 32, seed: int = 3, lo: int = 2048, hi: int = 65536) -> list:
    """The batched workload: `count` bonded tissue blocks of log-uniform size, each with its
    own domain, grid and seeded genome."""
    return [tissue(b=b, seed=seed * 1000 + k, dims=d) for k, d in enumerate(batch_sizes(count, lo, hi, seed))]


WORKLOADS = {
    "fluid_block_32k": lambda: fluid_block(32),
    "tissue_262k": lambda: tissue(64, 32),
}
