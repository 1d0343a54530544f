"""The reference solver's input and output files, for the hot-path configurations.

Reads what a wolfd2 run reads:

- the keyword deck (parse.f:lReadIPFile 32-878, lReadBCFile 1004-1951);
- the fluid property table (parse.f:lReadFldProps 2027-2249);
- the Plot3D grid (grid.f:ReadGrid 150-252);
- the unformatted restart file (file_manip.f:ReadRestart).

Writes what it writes: the restart file (file_manip.f:SaveRestart 364-401) and
the per-step PrintDiff table (string.f:PrintDiff 467-570).

The thermal keywords (energy equation, buoyancy, momentum/energy iterations,
filters, temperature walls and regions, heat generation) are read as the
reference reads them.  These features of the reference are outside the GPU
path, and a deck that turns one on is refused:

- small scales, trajectories;
- porous regions, whose coefficients come from the reference's permeability
  correlations.

The reference program linked against this library through the Fortran binding
serves those cases.
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from . import case

# parse.f:117-193 defaults of the keywords the hot path uses
DEFAULTS = dict(nts=0, mqiter=20, nPpeSolver=1, msorit=2000, qtol=1.0e-4, sortol=1.0e-8, sorrel=1.0,
                dlref=1.0, uref=1.0, tref=2.8357e2, tmax=1.0e3, dk=0.0,
                flPropFn="fluidprop.dat", flName="air01", resReadFl="restart.rst", resWriteFl="restart.rst",
                nrestart=0, nreswrite=0, nWrtRstFq=100, lPrntDiff=0, nPrDfFreq=1, lPrDfBann=0, nPrDfBnFq=20,
                cartesian=False, grid_file=None,
                nthermen=0, neqstate=0, nmeiter=1, dmeittol=1.0e-4, nfiltu=0, nfiltv=0, nfiltt=0,
                fpu=5.0e2, fpv=5.0e2, fpt=5.0)

# keywords that switch on physics outside the GPU path (parse.f:905-973)
REFUSED = {"small_scale", "trajectories"}
# keywords of outputs this runner does not produce, and the small-scale
# model's parameters (inert without small_scale); accepted and ignored
IGNORED = {"time_average_mode",
           "output_format", "output_prefix", "snapshots_prefix", "snapshots_freq", "snapshots_first",
           "time_series_prefix", "time_series_freq", "time_series_point", "atd_cu0", "atd_tscf", "atd_hscf",
           "atd_bnumcrit", "atd_rmapmax", "atd_rmapexp", "atd_temcf", "atd_ppe_solver", "atd_max_sor_iter",
           "atd_sor_tolerance", "atd_sor_relaxation", "atd_filter_u", "atd_filter_v", "atd_filter_t"}


class DeckError(ValueError):
    pass


@dataclass
class Deck:
    """Settings of one run (the keyword names of parse.f, values as parsed)."""
    settings: dict = field(default_factory=lambda: dict(DEFAULTS))
    nreg: tuple = (1, 1)
    i_borders: tuple = ()
    j_borders: tuple = ()
    bc: list = field(default_factory=list)   # (keyword, ireg, jreg, [side], words)
    profiles: list = field(default_factory=list)   # bc_profile lines: (ireg, jreg, side, var, words)


def _words(line: str) -> list:
    return line.split()


def read_input(path: str) -> Deck:
    """The input_parameters and boundary_conditions sections of an input deck."""
    d = Deck()
    s = d.settings
    section = None
    with open(path) as f:
        for ln, raw in enumerate(f, 1):
            w = _words(raw.split("#")[0])
            if not w:
                continue
            key = w[0].lower()
            if key == "section":
                section = w[1].lower() if len(w) > 1 else None
                continue
            if key == "end":
                section = None
                continue
            try:
                if section == "input_parameters":
                    _input_keyword(s, key, w[1:])
                elif section == "boundary_conditions":
                    _bc_keyword(d, key, w[1:])
                else:
                    raise DeckError(f"keyword '{key}' outside a section")
            except (IndexError, ValueError) as e:
                raise DeckError(f"{path}:{ln}: cannot parse '{raw.strip()}' ({e})") from None
    if not s["grid_file"]:
        raise DeckError("no grid_file in the deck")
    return d


def _input_keyword(s: dict, key: str, a: list) -> None:
    if key in REFUSED:
        raise DeckError(f"'{key}' is outside the GPU hot path (use the Fortran drop-in build for it)")
    if key in IGNORED:
        return
    if key == "thermal_energy":            # parse.f:334-340: the keyword alone switches it on
        s["nthermen"] = 1
    elif key == "buoyancy":
        s["neqstate"] = 1
    elif key == "max_me_iter":
        s["nmeiter"] = int(a[0])
    elif key == "me_tolerance":
        s["dmeittol"] = float(a[0])
    elif key in ("filter_u", "filter_v", "filter_t"):   # parse.f:404-438
        c = key[-1]
        s[f"nfilt{c}"] = 1
        s[f"fp{c}"] = float(a[0])
    elif key == "max_ref_temperature":
        s["tmax"] = float(a[0])
    elif key == "grid_file":
        s["grid_file"] = a[0]
        s["cartesian"] = len(a) > 1 and a[1].lower() == "cartesian_grid"
    elif key == "n_time_steps":
        s["nts"] = int(a[0])
    elif key == "time_step_size":
        s["dk"] = float(a[0])
    elif key == "max_ql_iter":
        s["mqiter"] = int(a[0])
    elif key == "ql_tolerance":
        s["qtol"] = float(a[0])
    elif key == "ppe_solver":
        if a[0].lower() not in case.PPE_SOLVERS:
            raise ValueError(f"unknown ppe_solver {a[0]}")
        s["nPpeSolver"] = case.PPE_SOLVERS[a[0].lower()]
    elif key == "max_sor_iter":
        s["msorit"] = int(a[0])
    elif key == "sor_tolerance":
        s["sortol"] = float(a[0])
    elif key == "sor_relaxation":
        s["sorrel"] = float(a[0])
    elif key == "ref_length":
        s["dlref"] = float(a[0])
    elif key == "ref_velocity":
        s["uref"] = float(a[0])
    elif key == "ref_temperature":
        s["tref"] = float(a[0])
    elif key == "fluid_prop_file":
        s["flPropFn"] = a[0]
    elif key == "fluid_prop_table":
        s["flName"] = a[0]
    elif key == "restart":
        s["nrestart"] = 1
        s["resReadFl"] = a[0]
    elif key == "write_restart":           # parse.f:679-700
        s["resWriteFl"] = a[0]
        if len(a) > 1:
            s["nreswrite"] = 2
            s["nWrtRstFq"] = int(a[1])
        else:
            s["nreswrite"] = 1
    elif key == "print_diff_freq":
        s["lPrntDiff"] = 1
        s["nPrDfFreq"] = int(a[0])
    elif key == "print_bann_freq":
        s["lPrDfBann"] = 1
        s["nPrDfBnFq"] = int(a[0])
    else:
        raise ValueError(f"unknown keyword '{key}'")


def _bc_keyword(d: Deck, key: str, a: list) -> None:
    if key == "number_of_regions":
        d.nreg = (int(a[0]), int(a[1]))
    elif key == "i_borders":
        d.i_borders = tuple(int(x) for x in a)
    elif key == "j_borders":
        d.j_borders = tuple(int(x) for x in a)
    elif key in ("wall", "inlet", "outlet"):
        d.bc.append((key, int(a[0]), int(a[1]), a[2].lower(), [x.lower() for x in a[3:]]))
    elif key == "blockage":
        d.bc.append((key, int(a[0]), int(a[1]), None, [x.lower() for x in a[2:]]))
    elif key in ("heat_generation_region", "fixed_temperature_region"):
        d.bc.append((key, int(a[0]), int(a[1]), None, [float(a[2])]))
    elif key in ("fixed_temperature_wall", "heat_flux_wall"):
        d.bc.append((key, int(a[0]), int(a[1]), a[2].lower(), [float(x) for x in a[3:4]]))
    elif key == "bc_profile":
        if a[2].lower() not in case.SIDES or a[3].lower() not in PROFILE_VARS or a[4].lower() not in PROFILE_SHAPES:
            raise ValueError("expected bc_profile IREG JREG SIDE u|v|p parabolic|sine|file ...")
        d.profiles.append((int(a[0]), int(a[1]), a[2].lower(), a[3].lower(), [a[4].lower()] + a[5:]))
    elif key == "porous_region":
        raise DeckError(f"'{key}' is outside the GPU hot path (use the Fortran drop-in build for it)")
    else:
        raise ValueError(f"unknown keyword '{key}'")


def _opt(words: list, names: tuple):
    """Value following the first of ``names`` in words, or None."""
    for k, w in enumerate(words):
        if w in names and k + 1 < len(words):
            return float(words[k + 1])
    return None


def read_fluid(path: str, name: str, tref: float) -> np.ndarray:
    """parse.f:lReadFldProps: the fluid's row (T, rho, cp, mu, k, Pr) at tref,
    interpolated linearly (utility.f:dLinInterp) between the two rows whose
    temperatures bracket it."""
    rows, found = [], False
    with open(path) as f:
        for raw in f:
            w = raw.split()
            if not w:
                continue
            if w[0].lower() == "fluid":
                found = w[1] == name
                rows = []
                continue
            if w[0].lower() == "end":
                if found:
                    break
                continue
            if found:
                rows.append([float(x) for x in w[:6]])
    if not found or not rows:
        raise DeckError(f"fluid '{name}' not found in {path}")
    if rows[0][0] > tref:
        raise DeckError(f"tref {tref} is below the first entry of fluid '{name}'")
    prev = rows[0]
    for cur in rows[1:]:
        if prev[0] <= tref <= cur[0]:
            out = [tref]
            for p in range(1, 6):
                y1, y2, x1, x2 = cur[p], prev[p], cur[0], prev[0]
                out.append(y1 + (y2 - y1) * (tref - x1) / (x2 - x1))
            return np.array(out)
        prev = cur
    raise DeckError(f"tref {tref} is outside the table of fluid '{name}'")


def read_grid(path: str):
    """grid.f:ReadGrid: 'nx ny' (or 'nx ny nz') then x, then y.  For a 3-D file
    every k-plane is read into the same 2-D array, so the last plane is kept."""
    with open(path) as f:
        head = f.readline().split()
        vals = np.array(f.read().split(), dtype=np.float64)
    if len(head) not in (2, 3):
        raise DeckError(f"syntax error in first line of grid file {path}")
    nx, ny = int(head[0]), int(head[1])
    nz = int(head[2]) if len(head) == 3 else 1
    n = nx * ny
    if vals.size < 2 * n * nz:
        raise DeckError(f"grid file {path} holds {vals.size} values, needs {2 * n * nz}")
    x = vals[(nz - 1) * n: nz * n].reshape(ny, nx).copy()
    y = vals[nz * n + (nz - 1) * n: nz * n + nz * n].reshape(ny, nx).copy()
    return x, y


def build_case(d: Deck, rho: float, nx: int, ny: int, x=None, y=None, case_dir: str = ".") -> case.Regions:
    """Region tables of the boundary_conditions section, with the reference's
    scalings: velocities / uref (parse.f:1532, 1630), pressures / (rho uref^2)
    (parse.f:1550, 1739).  bc_profile lines need the grid nodes x, y."""
    uref = d.settings["uref"]
    pscale = rho * uref * uref
    walls, inlets, outlets, blocks = [], [], [], []
    for key, ir, jr, side, w in d.bc:
        if key == "blockage":
            blocks.append((ir, jr))
            continue
        if key in THERMAL_ONLY:
            continue
        if side not in case.SIDES:
            raise DeckError(f"bad side '{side}' for {key} {ir} {jr}")
        press = _opt(w, ("press", "p"))
        pr = 0.0 if press is None else press / pscale
        if key == "wall":
            kind = "no_stress" if "no_stress" in w else "no_slip"
            tv = _opt(w, ("tangent_vel", "tv"))
            walls.append((ir, jr, side, kind, 0.0 if tv is None else tv / uref, pr))
        elif key == "inlet":
            nv = _opt(w, ("normal_vel", "nv"))
            tv = _opt(w, ("tangent_vel", "tv"))
            nv = 0.0 if nv is None else nv / uref
            tv = 0.0 if tv is None else tv / uref
            u, v = (nv, tv) if side in ("west", "east") else (tv, nv)
            inlets.append((ir, jr, side, u, v, 0.0))
        else:
            kind = "fully_dev" if ("fully_dev" in w or "fd" in w) else "mass_cons"
            outlets.append((ir, jr, side, kind, 0.0, 0.0, pr))
    # make_regions applies blockages, then walls, inlets and outlets; the
    # reference applies the lines in deck order, which differs only for a deck
    # that gives one region face two conditions
    reg = case.make_regions(nx, ny, nreg=d.nreg, i_borders=d.i_borders, j_borders=d.j_borders, walls=walls,
                            inlets=inlets, outlets=outlets, blockages=blocks)
    if d.profiles:
        if x is None or y is None:
            raise DeckError("bc_profile needs the grid nodes")
        for ir, jr, side, var, w in d.profiles:
            _apply_profile(reg, x, y, ir, jr, case.SIDES[side], PROFILE_VARS[var], w,
                           pscale if var == "p" else uref, case_dir)
    return reg


# ------------------------------------------------------------ thermal tables
THERMAL_ONLY = ("heat_generation_region", "fixed_temperature_region", "fixed_temperature_wall", "heat_flux_wall")


def build_thermal(d: Deck, reg: case.Regions, dconduct: float, pe: float) -> case.Thermal:
    """Thermal region tables of the boundary_conditions section, line by line
    in deck order as parse.f:lReadBCFile applies them, after the defaults of
    parse.f:InitBCFlags 2350-2375 (no source, internal sides, adiabatic outer
    walls).  Temperatures are scaled as (T - tref)/(tmax - tref) into
    dBCVal(.,.,side,4) of ``reg``; heat generation is made non-dimensional by
    bound_cond.f:NDHGenCoef 1995-2030."""
    s = d.settings
    tref, tmax, dlref = s["tref"], s["tmax"], s["dlref"]
    ni, nj = int(reg.nreg[0]), int(reg.nreg[1])
    th = case.make_thermal(reg)

    def scaled(x):
        return (x - tref) / (tmax - tref)

    def set_side(ir, jr, side, bt, value=None):
        th.nTemBdTp[ir - 1, jr - 1, side - 1] = bt
        if value is not None:
            reg.dBCVal[ir - 1, jr - 1, side - 1, case.VAR_T - 1] = value

    def neighbours(ir, jr):
        for di, dj, side in ((-1, 0, case.EAST), (1, 0, case.WEST), (0, -1, case.NORTH), (0, 1, case.SOUTH)):
            a, b = ir + di, jr + dj
            if 1 <= a <= ni and 1 <= b <= nj:
                yield a, b, side

    for key, ir, jr, side, w in d.bc:
        if key == "blockage":                       # parse.f:1318-1392
            for sd in (case.WEST, case.EAST, case.SOUTH, case.NORTH):
                set_side(ir, jr, sd, case.BT_HTFLUX)
            for a, b, sd in neighbours(ir, jr):
                set_side(a, b, sd, case.BT_HTFLUX)
            tv = _opt(w, ("temper", "t"))
            if tv is not None:
                dum = scaled(tv)
                th.nTRgType[ir - 1, jr - 1] = case.RT_TEMPER
                th.dTRgVal[ir - 1, jr - 1] = dum
                for sd in (case.WEST, case.EAST, case.SOUTH, case.NORTH):
                    set_side(ir, jr, sd, case.BT_TEMPER, dum)
                for a, b, sd in neighbours(ir, jr):
                    set_side(a, b, sd, case.BT_TEMPER, dum)
        elif key in ("wall", "inlet", "outlet"):   # parse.f:1553-1566, 1655-1668, 1738-1751
            tv = _opt(w, ("temper", "t"))
            if tv is not None:
                set_side(ir, jr, case.SIDES[side], case.BT_TEMPER, scaled(tv))
        elif key == "heat_generation_region":      # parse.f:1768-1800
            th.nTRgType[ir - 1, jr - 1] = case.RT_HEATGN
            th.dTRgVal[ir - 1, jr - 1] = w[0]
        elif key == "fixed_temperature_region":    # parse.f:1805-1845
            dum = scaled(w[0])
            th.nTRgType[ir - 1, jr - 1] = case.RT_TEMPER
            th.dTRgVal[ir - 1, jr - 1] = dum
            for sd in (case.WEST, case.EAST, case.SOUTH, case.NORTH):
                set_side(ir, jr, sd, case.BT_TEMPER, dum)
        elif key == "fixed_temperature_wall":      # parse.f:1847-1882
            if side not in case.SIDES or not w:
                raise DeckError(f"bad fixed_temperature_wall {ir} {jr}")
            set_side(ir, jr, case.SIDES[side], case.BT_TEMPER, scaled(w[0]))
        elif key == "heat_flux_wall":              # parse.f:1886-1926 (adiabatic)
            if side not in case.SIDES:
                raise DeckError(f"bad heat_flux_wall {ir} {jr}")
            set_side(ir, jr, case.SIDES[side], case.BT_HTFLUX)
    for jr in range(1, nj + 1):                    # bound_cond.f:NDHGenCoef
        for ir in range(1, ni + 1):
            if th.nTRgType[ir - 1, jr - 1] == case.RT_HEATGN:
                th.dHGSTval[ir - 1, jr - 1] = (th.dTRgVal[ir - 1, jr - 1] * (dlref * dlref)) / \
                    (pe * dconduct * (tmax - tref))
    return th


def read_gas_constant(path: str, name: str) -> float:
    """The gas constant of a fluid table entry, 'fluid NAME gasconstant R'
    (parse.f:2130-2145); 0 when the entry gives none."""
    with open(path) as f:
        for raw in f:
            w = raw.split()
            if len(w) >= 2 and w[0].lower() == "fluid" and w[1] == name:
                if len(w) >= 4 and w[2] == "gasconstant":
                    return float(w[3])
                return 0.0
    raise DeckError(f"fluid '{name}' not found in {path}")


# ------------------------------------------------------- boundary profiles
# An extension with no counterpart in the reference, whose boundary values are
# one constant per region side (dBCVal).  In the boundary_conditions section:
#
#   bc_profile IREG JREG SIDE VAR parabolic PEAK     4 PEAK s (1 - s)
#   bc_profile IREG JREG SIDE VAR sine A1 A2 ... An   sum_k Ak sin(k pi s)
#   bc_profile IREG JREG SIDE VAR file NAME           one value per grid line
#
# VAR is u, v or p; s is the position along the side (0 at its first border
# node, 1 at its last) of the point each grid line's value applies to
# (case.side_positions).  A file lists ny+2 values (west/east sides, j = 0..ny+1)
# or nx+2 (south/north, i = 0..nx+1).  Values are dimensional and scaled like
# the constants: velocities / uref, pressures / (rho uref^2).
PROFILE_VARS = {"u": case.VAR_U, "v": case.VAR_V, "p": case.VAR_P}
PROFILE_SHAPES = ("parabolic", "sine", "file")


def _apply_profile(reg, x, y, ir, jr, side, var, w, scale, case_dir):
    shape, args = w[0], w[1:]
    if shape == "file":
        with open(os.path.join(case_dir, args[0])) as f:
            vals = np.array(f.read().split(), dtype=np.float64)
    else:
        sp = case.side_positions(x, y, reg, ir, jr, side, var)
        if shape == "parabolic":
            vals = case.parabolic_profile(sp, float(args[0]))
        else:
            vals = case.sine_profile(sp, [float(a) for a in args])
    try:
        reg.set_profile(ir, jr, side, var, vals / scale)
    except ValueError as e:
        raise DeckError(f"bc_profile {ir} {jr}: {e}") from None


# ------------------------------------------------------------ restart files
def _records(buf: bytes):
    off = 0
    while off < len(buf):
        (n,) = struct.unpack_from("<i", buf, off)
        rec = buf[off + 4: off + 4 + n]
        (n2,) = struct.unpack_from("<i", buf, off + 4 + n)
        if n != n2:
            raise DeckError("bad record marker in restart file")
        yield rec
        off += 8 + n


RESTART_FIELDS = ("p", "u", "v", "t", "pss", "uss", "vss", "tss")


def read_restart(path: str) -> dict:
    """Unformatted restart (file_manip.f:387-398): k,dtime / nx,ny / p,u,v,t,pss,uss,vss,tss.

    Each field is returned with shape (ny+2, nx+2), element [j, i] = f(i, j).
    """
    with open(path, "rb") as f:
        recs = list(_records(f.read()))
    k, dtime = struct.unpack("<id", recs[0][:12])
    nx, ny = struct.unpack("<ii", recs[1][:8])
    data = np.frombuffer(recs[2], dtype="<f8")
    n = (nx + 2) * (ny + 2)
    out = {"k": k, "dtime": dtime, "nx": nx, "ny": ny}
    for m, name in enumerate(RESTART_FIELDS):
        out[name] = data[m * n:(m + 1) * n].reshape(ny + 2, nx + 2).copy()
    return out


def write_restart(path: str, k: int, dtime: float, nx: int, ny: int, fields: dict) -> None:
    """Inverse of read_restart; missing fields are written as zeros."""
    def rec(b: bytes) -> bytes:
        return struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))
    zeros = np.zeros((ny + 2, nx + 2))
    body = b"".join(np.ascontiguousarray(fields.get(nm, zeros), dtype="<f8").tobytes() for nm in RESTART_FIELDS)
    with open(path, "wb") as f:
        f.write(rec(struct.pack("<id", k, dtime)))
        f.write(rec(struct.pack("<ii", nx, ny)))
        f.write(rec(body))


# ------------------------------------------------------------ PrintDiff table
def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d editing (scale factor 0): 0.ddd...E+ee, right-justified."""
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = "Infinity" if x > 0 else "-Infinity"
    elif x == 0.0:
        s = ("-" if math.copysign(1.0, x) < 0 else "") + "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        ex = int(e) + 1
        es = f"E{'+' if ex >= 0 else '-'}{abs(ex):02d}" if abs(ex) <= 99 else f"{'+' if ex >= 0 else '-'}{abs(ex):03d}"
        s = ("-" if x < 0 else "") + "0." + digits + es
    if len(s) > w and s.startswith("0."):
        s = s[1:]
    if len(s) > w and s.startswith("-0."):
        s = "-" + s[2:]
    return s.rjust(w) if len(s) <= w else "*" * w


PRINTDIFF_RULE = " " + "-" * 40 + " " * 30      # a character*70 variable holding 40 dashes
PRINTDIFF_HEAD = " " + f"{'Tstp':>6}{'Time   ':>13}{'QlIt':>5}{'PpeIt':>7}{'|p1-pn|':>10}" + \
    "".join(f"{h:>13}" for h in ("|u1-un|", "|v1-vn|"))


# with the energy equation: a character*78 rule and four a12 norm headings
PRINTDIFF_RULE_T = " " + "-" * 40 + " " * 38
PRINTDIFF_HEAD_T = " " + f"{'Tstp':>6}{'Time   ':>13}{'QlIt':>5}{'PpeIt':>7}{'|p1-pn|':>10}" + \
    "".join(f"{h:>12}" for h in ("|u1-un|", "|v1-vn|", "|t1-tn|"))


def print_diff_row(k: int, time: float, nql: int, nsor: int, dif, thermal: bool = False) -> str:
    """string.f:PrintDiff row '(1x,i6,e13.6,i4,a1,i6,4(e13.6))' with three
    norms, or '(1x,i6,e13.6,i4,a1,i6,4(e12.5))' with four (energy equation)."""
    mark = "*" if nql < 0 else " "
    head = " " + f"{k:6d}" + fortran_e(time, 13, 6) + f"{nql:4d}" + mark + f"{nsor:6d}"
    if thermal:
        return head + "".join(fortran_e(float(v), 12, 5) for v in dif[:4])
    return head + "".join(fortran_e(float(v), 13, 6) for v in dif[:3])
