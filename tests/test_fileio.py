"""The reference's scene files (SimParams.txt, genome.csv, SavePointsCSV2 point files) against
the demo inputs the reference ships (tests/golden/demo/, copied by
tests/golden/copy_demo_inputs.py), and the binary checkpoint against the oracle."""
import os
from fractions import Fraction

import numpy as np
import pytest

from morphogenesis_amd import _ffi as F
from morphogenesis_amd import fileio as fio
from morphogenesis_amd import scenes

DEMO = os.path.join(os.path.dirname(__file__), "golden", "demo")
GENOME = os.path.join(DEMO, "genome.csv")
PARAMS = os.path.join(DEMO, "SimParams.txt")
POINTS = os.path.join(DEMO, "particles_pos_vel_color100001.csv")


def read(path):
    with open(path) as fh:
        return fh.read()


# ---------------------------------------------------------------- numbers
def exact_f32(s):
    """Nearest float32 by brute force over the two float32 values that bracket the decimal."""
    fr = Fraction(s)
    lo = np.float32(float(fr))
    cands = [np.nextafter(lo, np.float32(-np.inf)), lo, np.nextafter(lo, np.float32(np.inf))]
    errs = [abs(Fraction(float(c)) - fr) for c in cands]
    best = min(errs)
    ties = [c for c, e in zip(cands, errs) if e == best]
    return ties[0] if len(ties) == 1 else [c for c in ties if int(np.float32(c).view(np.uint32)) % 2 == 0][0]


def test_decimal_to_float32_rounds_once():
    one = np.float32(1.0)
    nxt = np.nextafter(one, np.float32(2))
    mid = Fraction(float(one)) + (Fraction(float(nxt)) - Fraction(float(one))) / 2
    above = mid + Fraction(1, 10 ** 30)              # rounds up; via double it is a tie -> 1.0
    s_above = _decimal(above, 40)
    assert float(s_above) == float(mid)             # the double is the midpoint itself
    assert fio.to_f32([s_above])[0] == nxt
    assert fio.to_f32([_decimal(mid, 40)])[0] == one  # exact tie: even mantissa
    rng = np.random.default_rng(1)
    toks = ["%.*f" % (int(rng.integers(1, 9)), x) for x in rng.normal(0, 1e3, 300)]
    toks += ["%.9e" % x for x in rng.lognormal(0, 20, 300)]
    got = fio.to_f32(toks)
    want = np.array([exact_f32(t) for t in toks], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.signbit(fio.to_f32(["-0.000000"])[0])
    assert np.isnan(fio.to_f32(["-nan"])[0]) and np.signbit(fio.to_f32(["-nan"])[0])


def _decimal(fr: Fraction, digits: int) -> str:
    """Exact decimal expansion of a positive fraction that has one within `digits` places."""
    ip = fr.numerator // fr.denominator
    frac = (fr - ip) * 10 ** digits
    assert frac.denominator == 1
    return f"{ip}.{int(frac):0{digits}d}"


def test_c_integer_conversions():
    assert fio.c_int("010") == 8 and fio.c_int("0x1F") == 31 and fio.c_int("-7") == -7
    assert fio.c_int("0") == 0 and fio.c_int("16") == 16
    assert fio.c_uint("-1") == 0xFFFFFFFF and fio.c_uint("4278203136") == 4278203136
    with pytest.raises(fio.FormatError):
        fio.c_int("1.5")
    assert fio.cf(np.float32(0.002)) == "0.002000" and fio.cf(np.float32(-np.nan)) in ("-nan", "nan")


# ---------------------------------------------------------------- genome
def test_demo_genome_is_the_example_genome():
    """The demo genome.csv was written from SetupExampleGenome; scenes.example_genome restates
    that function field by field, independently of the reader."""
    g = fio.read_genome(GENOME)
    ex = scenes.example_genome()
    assert bytes(g) == bytes(ex)
    assert list(g.secrete[0][:4]) == [1, 1, 3, 4] and g.secrete[0][2 * F.NUM_TF] == 2
    assert list(g.activate[0][:2]) == [5, 6] and g.activate[0][2 * F.NUM_GENES] == 1
    assert np.float32(g.param[1][6]) == np.float32(1e9) and np.float32(g.tanh_param[0][4]) == np.float32(1.008)


def test_write_genome_reproduces_the_demo_bytes(tmp_path):
    p = fio.write_genome(scenes.example_genome(), str(tmp_path))
    assert read(p) == read(GENOME)


def test_genome_round_trip_of_every_field(tmp_path):
    rng = np.random.default_rng(4)
    g = scenes.tissue_genome(3)
    for i in range(F.NUM_GENES):
        g.mutability[i] = int(rng.integers(0, 2 ** 32))      # prints negative through %i
        for j in range(2 * F.NUM_TF + 1):
            g.secrete[i][j] = int(rng.integers(-2 ** 31, 2 ** 31))
    for r in range(3):
        for k in range(12):
            g.param[r][k] = float(np.float32(rng.normal(0, 100)))
    p = fio.write_genome(g, str(tmp_path))
    h = fio.read_genome(p)
    a, b = np.frombuffer(bytes(g), np.uint32), np.frombuffer(bytes(h), np.uint32)
    # integers exact; floats as "%f" prints them
    nint = (F.NUM_GENES * (2 + F.NUM_GENES) + 2 * F.NUM_TF + F.NUM_GENES * (2 * F.NUM_TF + 1)
            + F.NUM_GENES * (2 * F.NUM_GENES + 1))
    assert np.array_equal(a[:nint], b[:nint])
    for r in range(3):
        for k in range(12):
            assert np.float32(h.param[r][k]) == fio.to_f32([fio.cf(np.float32(g.param[r][k]))])[0]
    assert read(fio.write_genome(h, str(tmp_path))) == read(p)


def test_genome_size_mismatch_is_an_error(tmp_path):
    bad = read(GENOME).replace("num genes = 16", "num genes = 12", 1)
    (tmp_path / "g.csv").write_text(bad)
    with pytest.raises(fio.FormatError):
        fio.read_genome(str(tmp_path / "g.csv"))


# ---------------------------------------------------------------- SimParams
def test_sim_params_read_and_write_reproduce_the_demo(tmp_path):
    sp = fio.read_sim_params(PARAMS)
    assert sp.time == 0.0 and np.float32(sp.dt) == np.float32(0.003)
    assert sp.vec["PVOLMAX"] == (10.0, 20.0, 50.0) and sp.vec["PPLANE_GRAV_DIR"] == (0.0, -1.0, 0.0)
    assert np.float32(sp.param["PMASS"]) == np.float32(0.000205)
    assert read(fio.write_sim_params(sp, str(tmp_path))) == read(PARAMS)


def test_sim_params_missing_entry_is_an_error(tmp_path):
    text = "\n".join(ln for ln in read(PARAMS).split("\n") if "PVISC" not in ln)
    (tmp_path / "SimParams.txt").write_text(text)
    with pytest.raises(fio.FormatError):
        fio.read_sim_params(str(tmp_path / "SimParams.txt"))


def test_fluid_setup_derives_the_demo_values():
    """SetupSpacing / SetupGrid from the demo SimParams: the derived PDIST and PRESTDENSITY
    print as the demo file holds them (it was written after set-up), the grid is
    (VOL extent / cell) with cell = 2*sr/density = 1, and PBOUND = VOL -/+ 2 cells."""
    sp = fio.read_sim_params(PARAMS)
    params, grid, sp2 = fio.fluid_setup(sp)
    assert fio.cf(np.float32(sp2.param["PDIST"])) == "1.149425"
    assert fio.cf(np.float32(sp2.param["PRESTDENSITY"])) == "0.000135"
    assert sp2.param["PGRIDSIZE"] == 1.0
    assert list(grid.res) == [10, 20, 50] and grid.total == 10000 and grid.srch == 3
    assert list(params.pboundmin) == [2.0, 2.0, 2.0] and list(params.pboundmax) == [8.0, 18.0, 48.0]
    assert list(params.pgravity) == [0.0, -2.0, 0.0]
    assert params.pground_slope == np.float32(0.1) and params.pdamp == np.float32(100.0)
    # the same set-up through the scene builder's own path
    ref = scenes.make_params((0, 0, 0), (10, 20, 50), spacing=1.0, mass=sp.param["PMASS"],
                             gslope=sp.param["PGROUND_SLOPE"])
    assert bytes(ref) == bytes(params)
    assert bytes(scenes.make_grid((0, 0, 0), (10, 20, 50))) == bytes(grid)


# ---------------------------------------------------------------- point files
def demo_bounds():
    _, _, sp2 = fio.fluid_setup(fio.read_sim_params(PARAMS))
    return (sp2.vec["PBOUNDMIN"], sp2.vec["PBOUNDMAX"]), sp2.param["PVEL_LIMIT"]


def test_demo_points_load_as_setup_add_volume_makes_them():
    """The demo point file is SetupAddVolumeMorphogenesis2's output; scenes.base_buffers
    restates that function, so every per-particle field but position and the gene flags
    must come out equal."""
    bounds, vl = demo_bounds()
    pts = fio.read_points_csv2(POINTS, bounds=bounds, vel_limit=vl)
    assert (pts.n, pts.b) == (20, 6)
    B = pts.buffers
    assert B[F.FPOS][0, :3].tolist() == fio.to_f32(["5.538457", "6.329686", "5.720038"]).tolist()
    base = scenes.base_buffers(20, 6, B[F.FPOS][:, :3])
    for k in (F.FPARTICLE_ID, F.FMASS_RADIUS, F.FNERVEIDX, F.FELASTIDX, F.FPARTICLEIDX,
              F.FVEL, F.FVEVAL, F.FAGE, F.FCONC):
        assert np.array_equal(B[k], base[k]), F.BUFFER_NAMES[k]
    assert np.all(B[F.FCOLOR] == 4278203136)
    assert np.all(B[F.FEPIGEN][:3] == scenes.INT_MAX) and np.all(B[F.FEPIGEN][3:] == 0)


def test_save_points_reproduces_the_demo_bytes(tmp_path):
    pts = fio.read_points_csv2(POINTS)
    p = fio.save_points_csv2(pts.buffers, str(tmp_path), 1)
    assert os.path.basename(p) == "particles_pos_vel_color100001.csv"
    assert read(p) == read(POINTS)


def test_points_round_trip_bonded_tissue(tmp_path):
    sc = scenes.tissue(n_side=5, b=8, seed=3)
    sc.buffers[F.FVEL][:, :3] = np.random.default_rng(1).normal(0, 0.2, (sc.n, 3))
    p1 = fio.save_points_csv2(sc.buffers, str(tmp_path), 7)
    text = read(p1)
    assert "FELASTIDX[8*9]" in text and "Particle_Idx[8*2]" in text and ", 65535," in text
    pts = fio.read_points_csv2(p1)
    assert (pts.n, pts.b) == (sc.n, 8)
    for k in (F.FPARTICLE_ID, F.FMASS_RADIUS, F.FNERVEIDX, F.FPARTICLEIDX, F.FEPIGEN, F.FAGE,
              F.FCOLOR):
        assert np.array_equal(pts.buffers[k], sc.buffers[k]), F.BUFFER_NAMES[k]
    for w in (0, 5, 8):
        assert np.array_equal(pts.buffers[F.FELASTIDX][:, :, w], sc.buffers[F.FELASTIDX][:, :, w])
    printed = lambda a: fio.to_f32([fio.cf(x) for x in np.ravel(a)]).reshape(np.shape(a))
    assert np.array_equal(pts.buffers[F.FPOS][:, :3], printed(sc.buffers[F.FPOS][:, :3]))
    assert np.array_equal(pts.buffers[F.FCONC], printed(sc.buffers[F.FCONC]))
    rest = sc.buffers[F.FELASTIDX][:, :, 2].view(np.float32)
    assert np.array_equal(pts.buffers[F.FELASTIDX][:, :, 2].view(np.float32), printed(rest))
    # what the reader gives back, the writer prints identically
    p2 = fio.save_points_csv2(pts.buffers, str(tmp_path / ".."), 7)
    assert read(p2) == text


def test_points_out_of_bounds_is_an_error(tmp_path):
    rows = read(POINTS).split("\n")
    assert "4.631190,3.644084" in rows[4]
    rows[4] = rows[4].replace("4.631190,3.644084", "4.631190,1.500000", 1)   # below PBOUNDMIN.y = 2
    (tmp_path / "p.csv").write_text("\n".join(rows))
    bounds, vl = demo_bounds()
    with pytest.raises(ValueError, match="out of bounds"):
        fio.read_points_csv2(str(tmp_path / "p.csv"), bounds=bounds, vel_limit=vl)
    fio.read_points_csv2(str(tmp_path / "p.csv"))     # no bounds given: loads


def test_points_bad_bond_index_is_an_error(tmp_path):
    rows = read(POINTS).split("\n")
    rows[1] = rows[1].replace("65535, 0.000000", "40, 0.000000", 1)
    (tmp_path / "p.csv").write_text("\n".join(rows))
    with pytest.raises(fio.FormatError, match="names particle 40"):
        fio.read_points_csv2(str(tmp_path / "p.csv"))


def test_empty_slot_value_by_particle_count():
    assert fio._auto_empty(20, None) == 65535
    assert fio._auto_empty(70000, None) == F.NONE
    with pytest.raises(ValueError):
        fio._auto_empty(70000, 65535)


def test_load_scene_from_the_three_files():
    sc = fio.load_scene(PARAMS, POINTS, GENOME)
    assert (sc.n, sc.b, sc.c) == (20, 6, 10000)
    assert np.float32(sc.dt) == np.float32(0.003)
    assert bytes(sc.genome) == bytes(scenes.example_genome())


# ---------------------------------------------------------------- checkpoint
def shuffled_ids(sc, seed):
    """Particle IDs in a random order (bond word 5 follows), so that ID order and creation
    order differ."""
    perm = np.random.default_rng(seed).permutation(sc.n).astype(np.uint32) + 1
    rec = sc.buffers[F.FELASTIDX]
    sc.buffers[F.FPARTICLE_ID][:] = perm
    has = rec[:, :, 0] != F.NONE
    rec[:, :, 5] = np.where(has, perm[np.where(has, rec[:, :, 0], 0)], F.NONE)
    return sc


@pytest.mark.parametrize("shuffle", [False, True])
def test_checkpoint_resumes_the_oracle_exactly(built, tmp_path, shuffle):
    """Download every buffer, upload into a fresh instance, physical sort (the sort of a
    sorted state is the identity and restores the point counts), step: the same bits as the
    run that never stopped. The incoming-bond sum runs in particle-ID order, which the
    state carries."""
    from oracle.mg_oracle import Oracle

    sc = scenes.tissue(n_side=6, b=8, seed=2)
    if shuffle:
        sc = shuffled_ids(sc, 5)
    a = Oracle(sc.n, sc.b)
    a.load_scene(sc)
    a.Run2PhysicalSort()
    a.step(5, sc.dt, sc.flags)
    counts = a.counts()
    path = fio.save_checkpoint(a, str(tmp_path / "ck"), genome=sc.genome, params=sc.params,
                               grid=sc.grid)
    a.step(5, sc.dt, sc.flags)
    n, b, bufs, params, grid, genome, _ = fio.load_checkpoint(path)
    assert (n, b) == (sc.n, sc.b)
    assert bytes(params) == bytes(sc.params) and bytes(genome) == bytes(sc.genome)
    o = Oracle(n, b)
    o.set_grid(grid)
    o.set_params(params)
    o.set_genome(genome)
    o.TransferToCL(bufs)
    o.Run2PhysicalSort()
    assert o.counts() == counts
    o.step(5, sc.dt, sc.flags)
    for k in fio.CHECKPOINT_BUFFERS:
        x, y = a.download(k), o.download(k)
        assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              y.view(np.uint32) if y.dtype == np.float32 else y), F.BUFFER_NAMES[k]


# ---------------------------------------------------------------- VTP
def test_vtp_holds_the_reference_arrays(tmp_path):
    import xml.etree.ElementTree as ET

    sc = scenes.tissue(n_side=4, b=6, seed=2)
    bufs = sc.buffers
    bufs[F.FEPIGEN][6:11] = 0
    bufs[F.FEPIGEN][9, :5] = 1
    bufs[F.FEPIGEN][8, 5:9] = 1
    bufs[F.FDENSITY][:] = np.linspace(0.0, 2.0, sc.n).astype(np.float32)
    bufs[F.FELASTIDX][3, 1, 2] = np.float32(0.0).view(np.uint32)       # a broken bond
    p = fio.save_points_vtp(bufs, str(tmp_path), 12, (0, 0, 0), (7, 8, 9))
    assert p.endswith("particles_pos_vel_color100012.vtp")
    piece = ET.parse(p).getroot().find("PolyData/Piece")
    n, b = sc.n, sc.b
    assert int(piece.get("NumberOfPoints")) == n + 8 and int(piece.get("NumberOfLines")) == n * b

    def arr(section, name=None):
        for a in piece.find(section).findall("DataArray"):
            if name is None or a.get("Name") == name:
                return np.array(a.text.split(), dtype=np.float64)
        raise KeyError(name)

    pts = arr("Points").reshape(-1, 3)
    assert np.array_equal(pts[:n].astype(np.float32), bufs[F.FPOS][:, :3])
    assert pts[n + 7].tolist() == [7, 8, 9] and pts[n].tolist() == [0, 0, 0]
    names = [a.get("Name") for a in piece.find("PointData").findall("DataArray")]
    assert names[:10] == ["FAGE", "FCLR", "FPARTICLE_ID", "FMASS_RADIUS", "FNERVEIDX", "FVEL", "FVEVAL",
                          "FFORCE", "FPRESS", "FDENSITY,(1/density,density)"]
    assert names[10:26] == [f"FCONC_{t}" for t in range(16)] and names[-1] == "F_TISSUE_TYPE"
    mr = arr("PointData", "FMASS_RADIUS").reshape(-1, 2)
    assert mr[0].tolist() == [int(bufs[F.FMASS_RADIUS][0]) & 0xFFFF, int(bufs[F.FMASS_RADIUS][0]) >> 16]
    dens = arr("PointData", "FDENSITY,(1/density,density)").reshape(-1, 2)
    assert dens[0].tolist() == [0.0, 0.0]
    assert np.float32(dens[10, 1]) == np.float32(1) / bufs[F.FDENSITY][10]
    assert np.array_equal(arr("PointData", "FCONC_3")[:n].astype(np.float32), bufs[F.FCONC][:, 3])
    tt = arr("PointData", "F_TISSUE_TYPE")
    assert tt[:5].tolist() == [9] * 5 and tt[5:9].tolist() == [8] * 4 and tt[n:].tolist() == [0] * 8
    conn = arr("Lines", "connectivity").reshape(-1, 2).astype(np.int64)
    partner = bufs[F.FELASTIDX][:, :, 0].astype(np.int64).ravel()
    own = np.repeat(np.arange(n), b)
    rest = bufs[F.FELASTIDX][:, :, 2].view(np.float32).ravel()
    expect = np.where((partner == F.NONE) | (rest < 1e-6), own, partner)
    assert np.array_equal(conn[:, 0], own) and np.array_equal(conn[:, 1], expect)
    assert conn[3 * b + 1, 1] == 3                                       # broken bond drawn to itself
    rl = arr("CellData", "restlength")
    assert len(rl) == (n + 8) + n * b and np.all(rl[:n + 8] == 0)
    assert np.array_equal(rl[n + 8:].astype(np.float32), rest)
