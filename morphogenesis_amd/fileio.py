"""The reference's scene files: SimParams.txt, genome.csv and the particle point files
particles_pos_vel_color%04d.csv, read into and written from this package's structures
(MgParams, MgGrid, MgGenome and the per-particle buffers in the reference layout), plus a
binary checkpoint that resumes a run exactly.

  SimParams.txt   WriteSimParams / ReadSimParams    (src/fileCL_IO.cpp:471-634)
  genome.csv      WriteGenome / ReadGenome          (src/fileCL_IO.cpp:71-246)
  point files     SavePointsCSV2 / ReadPointsCSV2   (src/fileCL_IO.cpp:250-469)

The writers produce the reference writers' bytes (C "%f" / "%u" / "%i" formatting of the
same fields in the same order). The readers accept what the reference's fscanf readers
accept for such files; whitespace, and in the point files empty comma fields, carry no
meaning. Decimal fields are converted to float32 with correct rounding, as strtof does.

Departures from the reference, each deliberate:
  * the point-file reader keeps FCONC with its particle (FCONC[i][tf]); the reference's host
    accessor getConc stores tf-major (src/fluid_system.h:574) while its kernels read
    particle-major (src/fluid_system_cuda.cu:765), which scatters a loaded file's
    concentrations across particles;
  * bond word 6 (stress_sq integrator) is taken from the file; the reference reads it as a
    float but copies the unread integer scratch word (src/fileCL_IO.cpp:431,
    src/fluid_system.cpp:958);
  * the empty-slot value 65535 (UINT_MAXSIZE, src/host_CL.cpp:14) of files with at most
    65535 particles is mapped to this library's 0xFFFFFFFF and back;
  * errors raise FormatError / ValueError instead of exit().
"""
from __future__ import annotations

import dataclasses
import os
import re
from fractions import Fraction

import numpy as np

from . import _ffi as F

REF_EMPTY = 65535          # UINT_MAXSIZE of the reference's CUDA path (src/host_CL.cpp:14)


class FormatError(ValueError):
    pass


# ---------------------------------------------------------------------------- numbers
def _f32_exact(tok: str) -> np.float32:
    """Nearest float32 to a decimal string, ties to even (exact rational arithmetic)."""
    fr = Fraction(tok.strip())
    f = np.float32(float(fr))
    best, err = None, None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        if not np.isfinite(c):
            continue
        e = abs(Fraction(float(c)) - fr)
        if err is None or e < err or (e == err and int(np.float32(c).view(np.uint32)) & 1 == 0):
            best, err = c, e
    return np.float32(best)


def to_f32(tokens) -> np.ndarray:
    """Decimal strings to float32 as strtof rounds them. The double conversion is correctly
    rounded, and rounding that double to float32 can only differ from direct rounding
    when the double sits exactly halfway between two float32 values; those few tokens
    take the exact path."""
    toks = [t.strip() for t in tokens]
    try:
        d = np.array([float(t) for t in toks], np.float64)
    except ValueError as e:
        raise FormatError(f"not a decimal number: {e}") from None
    with np.errstate(over="ignore", invalid="ignore"):
        f = d.astype(np.float32)
        fin = np.isfinite(f) & (f != 0)
        up = np.nextafter(f, np.float32(np.inf)).astype(np.float64)
        dn = np.nextafter(f, np.float32(-np.inf)).astype(np.float64)
        f64 = f.astype(np.float64)
        mid = fin & ((d == (f64 + up) * 0.5) | (d == (f64 + dn) * 0.5))
    for k in np.flatnonzero(mid):
        f[k] = _f32_exact(toks[k])
    # C prints a negative NaN as "-nan"; keep its sign
    for k in np.flatnonzero(np.isnan(d)):
        if toks[k].lstrip().startswith("-"):
            f[k] = -np.abs(f[k])
    return f


_CINT = re.compile(r"^([+-]?)(0[xX][0-9a-fA-F]+|0[0-7]*|[1-9][0-9]*)$")


def c_int(tok: str) -> int:
    """scanf %i: decimal, 0x-hexadecimal or 0-octal, with sign."""
    m = _CINT.match(tok.strip())
    if not m:
        raise FormatError(f"not an integer: {tok!r}")
    sign, body = m.groups()
    if body[:2].lower() == "0x":
        v = int(body, 16)
    elif body[0] == "0" and len(body) > 1:
        v = int(body, 8)
    else:
        v = int(body)
    return -v if sign == "-" else v


def c_uint(tok: str) -> int:
    """scanf %u: decimal with optional sign, wrapped to 32 bits as strtoul's result is stored."""
    t = tok.strip()
    if not re.match(r"^[+-]?[0-9]+$", t):
        raise FormatError(f"not an unsigned integer: {tok!r}")
    return int(t) & 0xFFFFFFFF


def cf(v) -> str:
    """C printf "%f" of a float promoted to double."""
    x = float(v)
    if x != x:
        return "-nan" if np.signbit(v) else "nan"
    return "%f" % x


def _ci(v) -> str:
    """C printf "%i" of a 32-bit value (unsigned fields print as their signed reading)."""
    v = int(v) & 0xFFFFFFFF
    return str(v - (1 << 32) if v >= (1 << 31) else v)


def _fields(line: str) -> list:
    return [t for t in (s.strip() for s in line.split(",")) if t]


# ---------------------------------------------------------------------------- SimParams
# Entries of SimParams.txt in the order WriteSimParams prints them (src/fileCL_IO.cpp:587-626):
# ("t", name) scalar time/dt, ("p", PNAME) m_Param, ("v", PNAME) m_Vec, ("c", text) comment.
SIM_LAYOUT = [("t", "m_Time"), ("t", "m_DT")] + [("p", n) for n in (
    "PSIMSCALE", "PGRID_DENSITY", "PVISC", "PSURFACE_TENSION", "PRESTDENSITY", "PSPACING",
    "PMASS", "PRADIUS", "PDIST", "PSMOOTHRADIUS", "PINTSTIFF", "PEXTSTIFF", "PEXTDAMP",
    "PACCEL_LIMIT", "PVEL_LIMIT", "PGRAV", "PGROUND_SLOPE", "PFORCE_MIN", "PFORCE_MAX",
    "PFORCE_FREQ")] + [("v", "PPLANE_GRAV_DIR"), ("c", "// Default sim config"),
                       ("p", "PGRIDSIZE"), ("v", "PVOLMIN"), ("v", "PVOLMAX"), ("v", "PINITMIN"),
                       ("v", "PINITMAX"), ("p", "PFORCE_MIN"), ("p", "PFORCE_FREQ"),
                       ("p", "PGROUND_SLOPE")]


@dataclasses.dataclass
class SimParams:
    """m_Time, m_DT and the m_Param / m_Vec entries that SimParams.txt carries, as float32."""
    time: float = 0.0
    dt: float = 0.0
    param: dict = dataclasses.field(default_factory=dict)   # "PSIMSCALE" -> float
    vec: dict = dataclasses.field(default_factory=dict)     # "PVOLMIN" -> (x, y, z)

    def copy(self) -> "SimParams":
        return SimParams(self.time, self.dt, dict(self.param), dict(self.vec))


_NUM = r"([-+]?(?:[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?|[0-9]+\.|nan|inf|infinity))"
_SIM_T = re.compile(r"^(m_Time|m_DT)\s*=\s*" + _NUM + r"\s*$", re.I)
_SIM_P = re.compile(r"^m_Param\s*\[\s*(\w+)\s*\]\s*=\s*" + _NUM + r"\s*$", re.I)
_SIM_V = re.compile(r"^m_Vec\s*\[\s*(\w+)\s*\]\s*\.Set\s*\(\s*" + _NUM + r"\s*,\s*" + _NUM +
                    r"\s*,\s*" + _NUM + r"\s*\)\s*$", re.I)


def read_sim_params(path: str) -> SimParams:
    """ReadSimParams (src/fileCL_IO.cpp:471-530): every entry of SIM_LAYOUT must be present;
    an entry that appears twice takes its later value, as the reference's second read of
    PFORCE_MIN / PFORCE_FREQ / PGROUND_SLOPE does."""
    with open(path, "r") as fh:
        text = fh.read()
    sp = SimParams()
    seen = set()
    for raw in text.split("\n"):
        line = raw.strip()
        if not line or line.startswith("//"):
            continue
        m = _SIM_T.match(line)
        if m:
            v = float(to_f32([m.group(2)])[0])
            seen.add(m.group(1))
            if m.group(1) == "m_Time":
                sp.time = v
            else:
                sp.dt = v
            continue
        m = _SIM_P.match(line)
        if m:
            sp.param[m.group(1)] = float(to_f32([m.group(2)])[0])
            continue
        m = _SIM_V.match(line)
        if m:
            sp.vec[m.group(1)] = tuple(float(x) for x in to_f32(m.groups()[1:]))
            continue
        raise FormatError(f"{path}: unrecognised SimParams line {raw!r}")
    need_p = {n for k, n in SIM_LAYOUT if k == "p"}
    need_v = {n for k, n in SIM_LAYOUT if k == "v"}
    missing = sorted((need_p - sp.param.keys()) | (need_v - sp.vec.keys()) |
                     ({"m_Time", "m_DT"} - seen))
    if missing:
        raise FormatError(f"{path}: SimParams read failure, missing {missing}")
    return sp


def write_sim_params(sp: SimParams, directory: str) -> str:
    """WriteSimParams (src/fileCL_IO.cpp:532-634): <directory>/SimParams.txt."""
    lines = []
    for kind, name in SIM_LAYOUT:
        if kind == "t":
            lines.append(f"{name} = {cf(np.float32(sp.time if name == 'm_Time' else sp.dt))}")
        elif kind == "p":
            lines.append(f"m_Param [ {name} ] = {cf(np.float32(sp.param[name]))}")
        elif kind == "v":
            x, y, z = (cf(np.float32(c)) for c in sp.vec[name])
            lines.append(f"m_Vec [ {name} ].Set ( {x}, {y}, {z} )")
        else:
            lines.append(name)
    path = os.path.join(directory, "SimParams.txt")
    with open(path, "w") as fh:
        fh.write("".join(" " + s + "\n" for s in lines) + " ")
    return path


def fluid_setup_args(sp: SimParams, actuation_factor: float = 0.0, actuation_period: float = 1.0):
    """The set-up ReadPointsCSV2 runs before it reads particles (src/fileCL_IO.cpp:351-366):
    PGRIDSIZE = 2*PSMOOTHRADIUS/PGRID_DENSITY, SetupSPH_Kernels, SetupSpacing
    (src/fluid_system.cpp:2521-2546), then the argument lists it hands to SetupGrid
    (src/fluid_system.cpp:2073-2131) and, through UpdateParams, to FluidParamCL
    (src/fluid_system.cpp:612-628), in float32 host arithmetic. Returns (SetupGrid arguments
    (vmin, vmax, sim_scale, cell_size, smooth_radius), FluidParamCL arguments, SimParams with
    the derived entries updated)."""
    f = np.float32
    out = sp.copy()
    P = out.param
    ss, sr = f(P["PSIMSCALE"]), f(P["PSMOOTHRADIUS"])
    gridsize = f(f(f(2) * sr) / f(P["PGRID_DENSITY"]))
    P["PGRIDSIZE"] = float(gridsize)
    mass, rest = f(P["PMASS"]), f(P["PRESTDENSITY"])
    with np.errstate(divide="ignore", invalid="ignore"):
        pdist = F.powf(f(mass / rest), f(f(1) / f(3)))
    if f(P["PSPACING"]) == 0:
        # spacing from density
        spacing = f(f(pdist * f(0.87)) / ss)
        P["PSPACING"] = float(spacing)
    else:
        # density from spacing
        pdist = f(f(f(P["PSPACING"]) * ss) / f(0.87))
        rest = f(mass / F.powf(pdist, 3))
    P["PDIST"] = float(pdist)
    P["PRESTDENSITY"] = float(rest)
    off = 2.0 * float(f(gridsize / ss))           # float quotient, doubled in double
    vmin = out.vec["PVOLMIN"]
    vmax = out.vec["PVOLMAX"]
    bmin = [float(f(float(f(v)) + off)) for v in vmin]
    bmax = [float(f(float(f(v)) - off)) for v in vmax]
    out.vec["PBOUNDMIN"] = tuple(bmin)
    out.vec["PBOUNDMAX"] = tuple(bmax)
    grid_args = (tuple(vmin), tuple(vmax), float(ss), float(gridsize), float(sr))
    gdir = out.vec["PPLANE_GRAV_DIR"]
    g = [float(f(f(c) * f(P["PGRAV"]))) for c in gdir]
    param_args = (
        float(ss), float(sr), P["PRADIUS"], P["PMASS"], float(rest), tuple(bmin), tuple(bmax),
        P["PEXTSTIFF"], P["PINTSTIFF"], P["PVISC"], P["PSURFACE_TENSION"], P["PEXTDAMP"],
        P["PFORCE_MIN"], P["PFORCE_MAX"], P["PFORCE_FREQ"], P["PGROUND_SLOPE"], g[0], g[1], g[2],
        P["PACCEL_LIMIT"], P["PVEL_LIMIT"], actuation_factor, actuation_period)
    return grid_args, param_args, out


def fluid_setup(sp: SimParams, actuation_factor: float = 0.0, actuation_period: float = 1.0):
    """fluid_setup_args, then SetupGrid and FluidParamCL through the library's host derivation
    (mg_derive_grid / mg_derive_params). Returns (MgParams, MgGrid, SimParams with the derived
    entries updated)."""
    grid_args, param_args, out = fluid_setup_args(sp, actuation_factor, actuation_period)
    return F.params_from_fluid_param(*param_args), F.grid_setup(*grid_args), out


# ---------------------------------------------------------------------------- genome
_REMODEL_COLUMNS = ("elongation_threshold,\telongation_factor,\tstrength_threshold,\t"
                    "strengthening_factor,\t\t    max_rest_length,\tmin_rest_length,\t"
                    "max_modulus,\t\t\tmin_modulus,\t\t    elastLim,\tdefault_rest_length,\t"
                    "default_modulus,\tdefault_damping")
_TANH_COLUMNS = ("l_a (y-shift),\t l_b (y-scaling),\t l_c (x-scaling),\t l_d (x-shift),\t\t "
                 "s_a (y-shift),\t s_b (y-scaling),\t s_c (x-scaling),\t s_d (x-shift)")
_TF_TITLE = "Transcription Factors (tf_difusibility,tf_breakdown_rate) "
_REMODEL_TITLE = "Remodelling parameters, rows : elastin,collagen,apatite"
_REMODEL_SUB = ("collumns : /*triggering bond parameter changes*/, /*triggering particle "
                "changes*/, /*initial values for new bonds*/")
_TANH_TITLE = "Bond remodelling m_FGenome.tanh_param[3][8] parameters, rows : elastin,collagen,apatite"
_TANH_SUB = "collumns : lengthen/shorten, strengthen/weaken"


def write_genome(g: F.MgGenome, directory: str) -> str:
    """WriteGenome (src/fileCL_IO.cpp:159-228): <directory>/genome.csv. The free-text
    description of the FGenome struct that the reference appends after the data is left
    out; ReadGenome never reads it."""
    NG, NT = F.NUM_GENES, F.NUM_TF
    out = [f"num genes = {NG},\tnum transcription factors = {NT} \n",
           "mutability, delay,\tsensitivity[NUM_GENES]" + ",\t" * NG + "secrete[NUM_TF][2]" +
           ",,\t" * NT + "secrete[2*NUM_TF], activate[NUM_GENES][2]" + ",,\t" * NG +
           "activate[2*NUM_GENES] \n"]
    for i in range(NG):
        row = [f"{_ci(g.mutability[i])},\t", f"{_ci(g.delay[i])},\t\t"]
        row += [f"{_ci(g.sensitivity[i][j])},\t" for j in range(NG)]
        row.append(" \t\t")
        row += [f"{_ci(g.secrete[i][2 * j])},{_ci(g.secrete[i][2 * j + 1])},\t" for j in range(NT)]
        row.append(f"\t\t{_ci(g.secrete[i][2 * NT])},\t\t")
        row += [f"{_ci(g.activate[i][2 * j])},{_ci(g.activate[i][2 * j + 1])},\t" for j in range(NG)]
        row.append(f"\t\t{_ci(g.activate[i][2 * NG])},\t\t")
        row.append(" \n")
        out.append("".join(row))
    out.append("\n" + _TF_TITLE + "\n")
    out.append("".join(f"\t{_ci(g.tf_diffusability[t])},\t{_ci(g.tf_breakdown_rate[t])},\t"
                       for t in range(NT)))
    out.append("\n\n" + _REMODEL_TITLE + "\n" + _REMODEL_SUB + "\n" + _REMODEL_COLUMNS + "\n")
    for r in range(3):
        out.append("".join(f"\t{cf(np.float32(g.param[r][k]))},\t" for k in range(12)) + "\n")
    out.append("\n\n" + _TANH_TITLE + "\n" + _TANH_SUB + "\n" + _TANH_COLUMNS + "\n")
    for r in range(3):
        out.append("".join(f"\t{cf(np.float32(g.tanh_param[r][k]))},\t" for k in range(8)) + "\n")
    path = os.path.join(directory, "genome.csv")
    with open(path, "w") as fh:
        fh.write("".join(out))
    return path


_GENOME_HEAD = re.compile(r"^\s*num genes\s*=\s*(\S+?)\s*,\s*num transcription factors\s*=\s*(\S+)\s*$")


def read_genome(path: str) -> F.MgGenome:
    """ReadGenome (src/fileCL_IO.cpp:71-157). Gene rows hold mutability, delay,
    sensitivity[16], 16 secrete pairs + count, 16 activate pairs + count (%i); then the
    transcription factor row (tf_diffusability, tf_breakdown_rate pairs, %u), the three
    remodelling rows of 12 floats and the three tanh rows of 8 floats. The gene and TF
    counts must be 16, the sizes of FGenome."""
    with open(path, "r") as fh:
        lines = fh.read().split("\n")
    m = _GENOME_HEAD.match(lines[0]) if lines else None
    if not m:
        raise FormatError(f"{path}: missing 'num genes = .., num transcription factors = ..' line")
    ng, nt = c_int(m.group(1)), c_int(m.group(2))
    NG, NT = F.NUM_GENES, F.NUM_TF
    if ng != NG or nt != NT:
        raise FormatError(f"{path}: {ng} genes / {nt} TFs, this build's FGenome holds {NG} / {NT}")
    g = F.MgGenome()
    per_gene = 2 + NG + 2 * NT + 1 + 2 * NG + 1
    for i in range(NG):
        if 2 + i >= len(lines):
            raise FormatError(f"{path}: gene {i} missing")
        t = _fields(lines[2 + i])
        if len(t) != per_gene:
            raise FormatError(f"{path}: gene {i} has {len(t)} fields, expected {per_gene}")
        v = [c_int(x) for x in t]
        g.mutability[i] = v[0] & 0xFFFFFFFF
        g.delay[i] = v[1] & 0xFFFFFFFF
        for j in range(NG):
            g.sensitivity[i][j] = v[2 + j] & 0xFFFFFFFF
        o = 2 + NG
        for j in range(2 * NT + 1):
            g.secrete[i][j] = _i32(v[o + j])
        o += 2 * NT + 1
        for j in range(2 * NG + 1):
            g.activate[i][j] = _i32(v[o + j])

    def after(title, skip):
        for k in range(2 + NG, len(lines)):
            if lines[k].strip() == title.strip():
                rows = [ln for ln in lines[k + 1 + skip:] if ln.strip()]
                return rows
        raise FormatError(f"{path}: section '{title.strip()}' not found")

    tf = _fields(after(_TF_TITLE, 0)[0])
    if len(tf) != 2 * NT:
        raise FormatError(f"{path}: transcription factor row has {len(tf)} fields, expected {2 * NT}")
    for t in range(NT):
        g.tf_diffusability[t] = c_uint(tf[2 * t])
        g.tf_breakdown_rate[t] = c_uint(tf[2 * t + 1])
    for title, skip, ncol, dst in ((_REMODEL_TITLE, 2, 12, g.param), (_TANH_TITLE, 2, 8, g.tanh_param)):
        rows = after(title, skip)
        for r in range(3):
            vals = _fields(rows[r]) if r < len(rows) else []
            if len(vals) != ncol:
                raise FormatError(f"{path}: '{title[:24]}' row {r} has {len(vals)} fields, expected {ncol}")
            for k, x in enumerate(to_f32(vals)):
                dst[r][k] = float(x)
    return g


def _i32(v: int) -> int:
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >= (1 << 31) else v


# ---------------------------------------------------------------------------- point files
def points_header(b: int) -> str:
    """Column-title line of SavePointsCSV2 (src/fileCL_IO.cpp:274-283) for B bond slots."""
    s = ["i,, x coord, y coord, z coord\t\t x vel, y vel, z vel\t\t age,  color\t\t "
         f"FELASTIDX[{b}*{F.DATA_PER_BOND}]"]
    s += [f",({a})[0]curIdx, [1]elastLim, [2]restLn, [3]modulus, [4]damping, [5]partID, "
          "[6]stress_sq integrator, [7]stress integrator, [8]change-type,,  " for a in range(b)]
    s.append("\t\tParticle_ID, mass, radius, FNERVEIDX,\t\t " + f"Particle_Idx[{b}*2]")
    s += [f"{a},,, " for a in range(b)]
    s.append(f"\t\tFCONC[{F.NUM_TF}] " + "".join(f"{t}, " for t in range(F.NUM_TF)))
    s.append(f"\t\tFEPIGEN[{F.NUM_GENES}] " + "".join(f"{t}, " for t in range(F.NUM_GENES)))
    return "".join(s) + "\n"


def _auto_empty(n: int, empty):
    if empty is None:
        return REF_EMPTY if n <= REF_EMPTY else F.NONE
    if empty != F.NONE and n > empty:
        raise ValueError(f"empty-slot value {empty} collides with particle indices (N = {n})")
    return int(empty)


def save_points_csv2(buffers: dict, directory: str, frame: int, empty=None) -> str:
    """SavePointsCSV2 (src/fileCL_IO.cpp:250-334): <directory>/particles_pos_vel_color%04d.csv
    with frame + 100000 in the name. One row per particle slot, stopping at the first
    particle ID 0 as the reference does. `empty` is the value written for empty bond
    slots: 65535 (the reference's) when N <= 65535, else 0xFFFFFFFF."""
    pos = np.asarray(buffers[F.FPOS], np.float32)
    n = pos.shape[0]
    rec = np.asarray(buffers[F.FELASTIDX], np.uint32)
    b = rec.shape[1] if rec.ndim == 3 else 0
    rec = rec.reshape(n, b, F.DATA_PER_BOND).copy()
    pidx = np.asarray(buffers[F.FPARTICLEIDX], np.uint32).reshape(n, b, 2).copy()
    e = _auto_empty(n, empty)
    if e != F.NONE:
        for w in (0, 5):
            rec[:, :, w][rec[:, :, w] == F.NONE] = e
        pidx[pidx == F.NONE] = e
    recf = rec.view(np.float32)
    vel = np.asarray(buffers[F.FVEL], np.float32)
    age = np.asarray(buffers[F.FAGE], np.uint32)
    clr = np.asarray(buffers[F.FCOLOR], np.uint32)
    pid = np.asarray(buffers[F.FPARTICLE_ID], np.uint32)
    mr = np.asarray(buffers[F.FMASS_RADIUS], np.uint32)
    nerve = np.asarray(buffers[F.FNERVEIDX], np.uint32)
    conc = np.asarray(buffers[F.FCONC], np.float32).reshape(n, F.NUM_TF)
    epi = np.asarray(buffers[F.FEPIGEN], np.uint32).reshape(F.NUM_GENES, n)
    rows = [points_header(b)]
    for i in range(n):
        if pid[i] == 0:
            break
        r = [f"{i},,{cf(pos[i, 0])},{cf(pos[i, 1])},{cf(pos[i, 2])},\t"
             f"{cf(vel[i, 0])},{cf(vel[i, 1])},{cf(vel[i, 2])},\t {age[i]}, {clr[i]},, \t"]
        for a in range(b):
            w, wf = rec[i, a], recf[i, a]
            r.append(f"{w[0]}, {cf(wf[1])}, {cf(wf[2])}, {cf(wf[3])}, {cf(wf[4])}, {w[5]}, "
                     f"{cf(wf[6])}, {cf(wf[7])}, {w[8]}, \t\t")
        m = int(mr[i])
        r.append(f" \t{pid[i]}, {m & 0xFFFF}, {m >> 16}, {nerve[i]}, \t\t")
        r.append("".join(f"{pidx[i, a, 0]}, {pidx[i, a, 1]},, " for a in range(b)) + "\t\t")
        r.append("".join(f"{cf(c)}, " for c in conc[i]) + "\t\t")
        r.append("".join(f"{epi[g, i]}, " for g in range(F.NUM_GENES)) + " \n")
        rows.append("".join(r))
    path = os.path.join(directory, "particles_pos_vel_color%04d.csv" % (frame + 100000))
    with open(path, "w") as fh:
        fh.write("".join(rows))
    return path


_HDR_BONDS = re.compile(r"FELASTIDX\[\s*(\d+)\s*\*\s*(\d+)\s*\]")
_HDR_PIDX = re.compile(r"Particle_Idx\[\s*(\d+)\s*\*\s*2\s*\]")
_HDR_CONC = re.compile(r"FCONC\[\s*(\d+)\s*\]")
_HDR_EPI = re.compile(r"FEPIGEN\[\s*(\d+)\s*\]")


@dataclasses.dataclass
class Points:
    n: int
    b: int
    buffers: dict


def read_points_csv2(path: str, bounds=None, vel_limit: float | None = None, empty=None) -> Points:
    """ReadPointsCSV2 (src/fileCL_IO.cpp:336-469) + AddParticleMorphogenesis2
    (src/fluid_system.cpp:930-984): one particle per row after the column-title line, which
    gives B (FELASTIDX[B*9]). Velocity is loaded, FVEVAL / FFORCE / FPRESS / FDENSITY start at
    zero. With `bounds` = (PBOUNDMIN, PBOUNDMAX) and `vel_limit`, a particle outside the
    bounds or faster than the limit is an error, as in the reference."""
    with open(path, "r") as fh:
        text = fh.read()
    lines = text.split("\n")
    head = lines[0]
    mb, mp, mc, me = (r.search(head) for r in (_HDR_BONDS, _HDR_PIDX, _HDR_CONC, _HDR_EPI))
    if not (mb and mp and mc and me):
        raise FormatError(f"{path}: not a SavePointsCSV2 file (column-title line)")
    b, dpb = int(mb.group(1)), int(mb.group(2))
    if dpb != F.DATA_PER_BOND or int(mp.group(1)) != b:
        raise FormatError(f"{path}: bond layout {b}*{dpb} / Particle_Idx[{mp.group(1)}*2]")
    if int(mc.group(1)) != F.NUM_TF or int(me.group(1)) != F.NUM_GENES:
        raise FormatError(f"{path}: FCONC[{mc.group(1)}] / FEPIGEN[{me.group(1)}], expected 16 / 16")
    body = [ln for ln in lines[1:] if ln.strip()]
    n = len(body)
    width = 9 + F.DATA_PER_BOND * b + 4 + 2 * b + F.NUM_TF + F.NUM_GENES
    toks = []
    for k, ln in enumerate(body):
        t = _fields(ln)
        if len(t) != width:
            raise FormatError(f"{path}: row {k + 1} has {len(t)} fields, expected {width}")
        toks.append(t)
    T = np.array(toks, dtype=object).reshape(n, width) if n else np.zeros((0, width), object)

    def ucols(lo, hi):
        return np.array([[c_uint(x) for x in row] for row in T[:, lo:hi]], np.uint32).reshape(n, hi - lo)

    def fcols(lo, hi):
        return to_f32(T[:, lo:hi].ravel().tolist()).reshape(n, hi - lo)

    bufs = {}
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    pos[:, :3] = fcols(1, 4)
    vel[:, :3] = fcols(4, 7)
    ac = ucols(7, 9)
    o = 9
    rec = np.zeros((n, b, F.DATA_PER_BOND), np.uint32)
    if b:
        words = T[:, o:o + F.DATA_PER_BOND * b].reshape(n, b, F.DATA_PER_BOND)
        for w in range(F.DATA_PER_BOND):
            col = words[:, :, w].ravel().tolist()
            if w in (0, 5, 8):
                rec[:, :, w] = np.array([c_uint(x) for x in col], np.uint32).reshape(n, b)
            else:
                rec[:, :, w] = to_f32(col).reshape(n, b).view(np.uint32)
    o += F.DATA_PER_BOND * b
    idm = ucols(o, o + 4)
    o += 4
    pidx = ucols(o, o + 2 * b).reshape(n, b, 2)
    o += 2 * b
    conc = fcols(o, o + F.NUM_TF)
    o += F.NUM_TF
    epi = ucols(o, o + F.NUM_GENES)
    e = _auto_empty(n, empty)
    if e != F.NONE:
        for w in (0, 5):
            rec[:, :, w][rec[:, :, w] == e] = F.NONE
        pidx[pidx == e] = F.NONE
    bad = (rec[:, :, 0] != F.NONE) & (rec[:, :, 0] >= n)
    if bad.any():
        i, a = np.argwhere(bad)[0]
        raise FormatError(f"{path}: particle {i} bond {a} names particle {rec[i, a, 0]} of {n}")
    if bounds is not None:
        lo = np.asarray(bounds[0], np.float32)
        hi = np.asarray(bounds[1], np.float32)
        v = vel[:, :3]
        v2 = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
        vl = np.float32(vel_limit) if vel_limit is not None else np.float32(np.inf)
        out = (pos[:, :3] < lo).any(1) | (pos[:, :3] > hi).any(1) | (v2 > vl * vl)
        if out.any():
            i = int(np.flatnonzero(out)[0])
            raise ValueError(f"{path}: particle out of bounds, row {i + 1}: pos {pos[i, :3].tolist()}, "
                             f"PBOUNDMIN {lo.tolist()}, PBOUNDMAX {hi.tolist()}, |v| limit {float(vl)}")
    bufs[F.FPOS] = pos
    bufs[F.FVEL] = vel
    bufs[F.FVEVAL] = np.zeros((n, 4), np.float32)
    bufs[F.FFORCE] = np.zeros((n, 4), np.float32)
    bufs[F.FPRESS] = np.zeros(n, np.float32)
    bufs[F.FDENSITY] = np.zeros(n, np.float32)
    bufs[F.FAGE] = ac[:, 0].copy()
    bufs[F.FCOLOR] = ac[:, 1].copy()
    bufs[F.FGCELL] = np.full(n, F.GRID_UNDEF, np.uint32)
    bufs[F.FGNDX] = np.zeros(n, np.uint32)
    bufs[F.FBIN] = np.zeros(n, np.uint32)
    bufs[F.FELASTIDX] = rec
    bufs[F.FPARTICLEIDX] = pidx
    bufs[F.FPARTICLE_ID] = idm[:, 0].copy()
    bufs[F.FMASS_RADIUS] = ((idm[:, 1] & 0xFFFF) | (idm[:, 2] << 16)).astype(np.uint32)
    bufs[F.FNERVEIDX] = idm[:, 3].copy()
    bufs[F.FCONC] = conc.astype(np.float32)
    bufs[F.FEPIGEN] = np.ascontiguousarray(epi.T)
    return Points(n, b, bufs)


# ---------------------------------------------------------------------------- VTP
def _vtk_array(name, vtype, ncomp, values, fmt) -> str:
    body = " ".join(fmt(v) for v in np.ravel(values))
    comp = f' NumberOfComponents="{ncomp}"' if ncomp > 1 else ""
    nm = f' Name="{name}"' if name else ""
    return f'        <DataArray type="{vtype}"{nm}{comp} format="ascii">{body}</DataArray>\n'


def save_points_vtp(buffers: dict, directory: str, frame: int, vol_min, vol_max) -> str:
    """SavePointsVTP2 (src/fileCL_IO.cpp:1223-1684): a VTK XML PolyData file
    <directory>/particles_pos_vel_color%04d.vtp (frame + 100000) with the particles up to the
    first one whose ID exceeds the particle count, the 8 corners of [vol_min, vol_max], and the
    reference's point arrays (FAGE, FCLR, FPARTICLE_ID, FMASS_RADIUS, FNERVEIDX, FVEL, FVEVAL,
    FFORCE, FPRESS, FDENSITY as (1/density, density), FCONC_t, FEPIGEN_g, F_TISSUE_TYPE; zeros for
    the corners) and per-bond arrays. Unlike the reference's writer, whose line insertion is
    commented out and which sets no vertices, every point gets a vertex cell and every bond slot a
    line cell (to the partner, or to the particle itself when the slot is empty or broken), so the
    per-bond arrays are cell data of a well-formed file (zeros on the vertex cells)."""
    pos = np.asarray(buffers[F.FPOS], np.float32)
    n = pos.shape[0]
    rec = np.asarray(buffers[F.FELASTIDX], np.uint32)
    b = rec.shape[1] if rec.ndim == 3 else 0
    rec = rec.reshape(n, b, F.DATA_PER_BOND)
    pid = np.asarray(buffers[F.FPARTICLE_ID], np.uint32)
    na = 0
    while na < n and pid[na] <= n:
        na += 1
    corners = np.array([[vol_max[0] if c & 1 else vol_min[0], vol_max[1] if c & 2 else vol_min[1],
                         vol_max[2] if c & 4 else vol_min[2]] for c in range(8)], np.float32)
    npts = na + 8

    def pad(a, ncomp=1):
        a = np.asarray(a)[:na].reshape(na, ncomp) if ncomp > 1 else np.asarray(a)[:na]
        z = np.zeros((8, ncomp) if ncomp > 1 else 8, a.dtype)
        return np.concatenate([a, z])

    fl = lambda v: str(np.float32(v))        # shortest string that reads back as the float32
    ui = lambda v: str(int(v))
    epi = np.asarray(buffers[F.FEPIGEN], np.uint32).reshape(F.NUM_GENES, n)
    conc = np.asarray(buffers[F.FCONC], np.float32).reshape(n, F.NUM_TF)
    inv = np.asarray(buffers[F.FDENSITY], np.float32)[:na]
    with np.errstate(divide="ignore"):
        dens = np.where(inv == 0, np.float32(0), np.float32(1) / np.where(inv == 0, np.float32(1), inv))
    mr = np.asarray(buffers[F.FMASS_RADIUS], np.uint32)
    tissue = np.select([epi[9] > 0, epi[6] > 0, epi[7] > 0, epi[10] > 0, epi[8] > 0],
                       [9, 6, 7, 10, 8], 0).astype(np.uint32)
    point_arrays = [
        ("FAGE", "UInt32", 1, pad(buffers[F.FAGE]), ui),
        ("FCLR", "UInt32", 1, pad(buffers[F.FCOLOR]), ui),
        ("FPARTICLE_ID", "UInt32", 1, pad(pid), ui),
        ("FMASS_RADIUS", "UInt32", 2, pad(np.stack([mr & 0xFFFF, mr >> 16], 1), 2), ui),
        ("FNERVEIDX", "UInt32", 1, pad(buffers[F.FNERVEIDX]), ui),
        ("FVEL", "Float32", 3, pad(np.asarray(buffers[F.FVEL])[:, :3], 3), fl),
        ("FVEVAL", "Float32", 3, pad(np.asarray(buffers[F.FVEVAL])[:, :3], 3), fl),
        ("FFORCE", "Float32", 3, pad(np.asarray(buffers[F.FFORCE])[:, :3], 3), fl),
        ("FPRESS", "Float32", 1, pad(buffers[F.FPRESS]), fl),
        ("FDENSITY,(1/density,density)", "Float32", 2, pad(np.stack([inv, dens], 1), 2), fl),
    ]
    point_arrays += [(f"FCONC_{t}", "Float32", 1, pad(conc[:, t]), fl) for t in range(F.NUM_TF)]
    point_arrays += [(f"FEPIGEN_{g}", "UInt32", 1, pad(epi[g]), ui) for g in range(F.NUM_GENES)]
    point_arrays.append(("F_TISSUE_TYPE", "UInt32", 1, pad(tissue), ui))
    # bond lines: (i, partner), self when empty / broken (rest length < 1e-6)
    recf = rec.view(np.float32)
    src = np.repeat(np.arange(na, dtype=np.int64), b)
    dst = rec[:na, :, 0].astype(np.int64).ravel()
    rest = recf[:na, :, 2].ravel()
    bad = (rest < 1e-6) | (dst >= na) | np.isnan(rest)
    dst = np.where(bad, src, dst)
    nl = na * b
    cz = lambda a: np.concatenate([np.zeros((npts,) + np.shape(a)[1:], np.asarray(a).dtype), a])
    uint6 = np.stack([src, np.repeat(pid[:na], b), rec[:na, :, 0].ravel(), rec[:na, :, 5].ravel(),
                      np.tile(np.arange(b), na), np.zeros(nl, np.int64)], 1).astype(np.uint32)
    cell_arrays = [("parent_idx, parent_ID, curr_idx, particle ID, bond index", "UInt32", 6, cz(uint6), ui)]
    for name, w in (("elastic_limit", 1), ("restlength", 2), ("modulus", 3), ("damping_coeff", 4),
                    ("strain_integrator", 7), ("strain_sq_integrator", 6)):
        cell_arrays.append((name, "Float32", 1, cz(recf[:na, :, w].ravel()), fl))
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="PolyData" version="0.1" byte_order="LittleEndian">\n', "  <PolyData>\n",
           f'    <Piece NumberOfPoints="{npts}" NumberOfVerts="{npts}" NumberOfLines="{nl}" '
           'NumberOfStrips="0" NumberOfPolys="0">\n', "      <PointData>\n"]
    out += [_vtk_array(*a) for a in point_arrays]
    out.append("      </PointData>\n      <CellData>\n")
    out += [_vtk_array(*a) for a in cell_arrays]
    out.append("      </CellData>\n      <Points>\n")
    out.append(_vtk_array("", "Float32", 3, np.concatenate([pos[:na, :3], corners]), fl))
    out.append("      </Points>\n      <Verts>\n")
    out.append(_vtk_array("connectivity", "Int64", 1, np.arange(npts), ui))
    out.append(_vtk_array("offsets", "Int64", 1, np.arange(1, npts + 1), ui))
    out.append("      </Verts>\n      <Lines>\n")
    out.append(_vtk_array("connectivity", "Int64", 1, np.stack([src, dst], 1), ui))
    out.append(_vtk_array("offsets", "Int64", 1, 2 * np.arange(1, nl + 1), ui))
    out.append("      </Lines>\n    </Piece>\n  </PolyData>\n</VTKFile>\n")
    path = os.path.join(directory, "particles_pos_vel_color%04d.vtp" % (frame + 100000))
    with open(path, "w") as fh:
        fh.write("".join(out))
    return path


# ---------------------------------------------------------------------------- scenes
def load_scene(sim_params_path: str, points_path: str, genome_path: str | None = None,
               flags: int = 0, name: str = "reference_files", check_bounds: bool = True):
    """A Scene from the reference's three input files, set up as ReadSimParams +
    ReadGenome + ReadPointsCSV2 leave a FluidSystem: dt is m_DT, the genome is the file's
    (or all zero with delay 1 and breakdown 1 when none is given)."""
    from . import scenes

    sp = read_sim_params(sim_params_path)
    params, grid, sp2 = fluid_setup(sp)
    bounds = (sp2.vec["PBOUNDMIN"], sp2.vec["PBOUNDMAX"]) if check_bounds else None
    pts = read_points_csv2(points_path, bounds=bounds, vel_limit=sp2.param["PVEL_LIMIT"])
    genome = read_genome(genome_path) if genome_path else scenes.empty_genome()
    return scenes.Scene(name, pts.n, pts.b, params, grid, genome, pts.buffers, float(sp.dt),
                        1, flags, {"sim_params": sp2, "time": sp.time})


# ---------------------------------------------------------------------------- checkpoints
CHECKPOINT_BUFFERS = F.PARTICLE_BUFFERS + F.GRID_BUFFERS


def save_checkpoint(sim, path: str, time: float | None = None, genome: F.MgGenome | None = None,
                    params: F.MgParams | None = None, grid: F.MgGrid | None = None) -> str:
    """Every buffer of `sim` (TransferFromCL of the particle and grid buffers) with the
    parameters, grid and genome, in one uncompressed .npz without pickled objects. Loading
    it into a fresh system and stepping reproduces the uninterrupted run bit for bit; the
    point-file format cannot (it prints 6 decimals and drops FVEVAL, FFORCE, FPRESS and
    FDENSITY)."""
    arrays = {F.BUFFER_NAMES[k]: sim.download(k) for k in CHECKPOINT_BUFFERS}
    for key, obj in (("params", params), ("grid", grid), ("genome", genome)):
        if obj is not None:
            arrays[f"struct_{key}"] = np.frombuffer(bytes(obj), np.uint8)
    arrays["meta"] = np.array([sim.n, sim.b, sim.c], np.int64)
    arrays["time"] = np.array([sim.m_Time if time is None else time], np.float64)
    np.savez(path, **arrays)
    return path if path.endswith(".npz") else path + ".npz"


def load_checkpoint(path: str):
    """(n, b, buffers, params, grid, genome, time) from save_checkpoint; structs that were not
    saved come back as None."""
    with np.load(path, allow_pickle=False) as z:
        n, b, c = (int(x) for x in z["meta"])
        names = {v: k for k, v in F.BUFFER_NAMES.items()}
        bufs = {names[k]: z[k] for k in z.files if k in names}
        out = {}
        for key, cls in (("params", F.MgParams), ("grid", F.MgGrid), ("genome", F.MgGenome)):
            k = f"struct_{key}"
            out[key] = cls.from_buffer_copy(z[k].tobytes()) if k in z.files else None
        time = float(z["time"][0])
    return n, b, bufs, out["params"], out["grid"], out["genome"], time
