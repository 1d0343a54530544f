"""Case set-up for the wolfd2 hot path: grid metrics, region tables, scalings.

This is the part of the reference's initialisation that produces the inputs of
the per-timestep calls (main.f:376-601): the metric arrays of grid.f, the
region/boundary tables of bound_cond.f:SetUpBCs and parse.f:InitBCFlags, and
the non-dimensional parameters of file_manip.f:ReadParam.  It runs once per
case on the host, in numpy, with the reference's operation order so that the
arrays are bit-identical to the reference's.

Array convention (the reference's): a field dimensioned (0:mnx,0:mny) in
column-major order is a C-contiguous numpy array of shape (mny+1, mnx+1)
indexed [j, i]; the leading dimension is ld = mnx+1.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# wolfd2.h:15-42
RM_BLOCKG, RM_INTERN, RM_POROUS = 0, 1, 2
BM_INTERN, BM_WALL1, BM_WALL2, BM_INLET, BM_OUTLT1, BM_OUTLT2 = 0, 1, 2, 3, 4, 5
WEST, EAST, SOUTH, NORTH = 1, 2, 3, 4
VAR_U, VAR_V, VAR_P, VAR_T = 1, 2, 3, 4
SIDES = {"west": WEST, "east": EAST, "south": SOUTH, "north": NORTH}

MGRI, MGRJ = 20, 10          # config.f:21-22

METRIC_NAMES = (
    "rau", "rbu", "rbv", "rgv", "ran", "rbn", "rgn", "rac", "rbc", "rgc",
    "dju", "djv", "djc", "djn", "xen", "yen", "xzn", "yzn", "xec", "yec",
    "xzc", "yzc", "xeu", "yeu", "xzv", "yzv", "xzu", "yzu", "xev", "yev",
)


def field_array(ld: int, ncol: int, dtype=np.float64) -> np.ndarray:
    """Zero-filled (0:ld-1, 0:ncol-1) Fortran field as a [j, i] numpy array."""
    return np.zeros((ncol, ld), dtype=dtype)


@dataclass
class Grid:
    """Nodes and metric arrays of grid.f:Grid 31-127.

    ``x``, ``y`` are the node coordinates on i=1..nx, j=1..ny as (ny, nx)
    arrays.  Metric arrays are (mny+1, mnx+1) with ld = mnx+1 and are zero
    outside i=1..nx, j=1..ny, exactly as the reference leaves them.
    """
    nx: int
    ny: int
    ld: int
    ncol: int
    metrics: dict = field(default_factory=dict)
    gx: np.ndarray | None = None
    gy: np.ndarray | None = None

    def __getattr__(self, name):
        m = self.__dict__.get("metrics")
        if m is not None and name in m:
            return m[name]
        raise AttributeError(name)


def compute_grid(x: np.ndarray, y: np.ndarray, dlref: float = 1.0,
                 mnx: int | None = None, mny: int | None = None) -> Grid:
    """grid.f:Grid without the file read: scale, MirrorPts, FullGrid, Metric."""
    ny, nx = x.shape
    mnx = nx + 1 if mnx is None else mnx
    mny = ny + 1 if mny is None else mny
    if nx + 1 > mnx or ny + 1 > mny:
        raise ValueError(f"grid {nx}x{ny} needs mnx>={nx + 1}, mny>={ny + 1} (grid.f:CheckGridSize)")
    ld, ncol = mnx + 1, mny + 1
    gx = field_array(ld, ncol)
    gy = field_array(ld, ncol)
    gx[1:ny + 1, 1:nx + 1] = x
    gy[1:ny + 1, 1:nx + 1] = y
    # grid.f:100-106
    gx[1:ny + 1, 1:nx + 1] = gx[1:ny + 1, 1:nx + 1] / dlref
    gy[1:ny + 1, 1:nx + 1] = gy[1:ny + 1, 1:nx + 1] / dlref

    # grid.f:MirrorPts 258-306 (arrays indexed [j, i])
    for a in (gx, gy):
        a[0, 1:nx + 1] = 2.0 * a[1, 1:nx + 1] - a[2, 1:nx + 1]
        a[ny + 1, 1:nx + 1] = 2.0 * a[ny, 1:nx + 1] - a[ny - 1, 1:nx + 1]
    for a in (gx, gy):
        a[1:ny + 1, 0] = 2.0 * a[1:ny + 1, 1] - a[1:ny + 1, 2]
        a[1:ny + 1, nx + 1] = 2.0 * a[1:ny + 1, nx] - a[1:ny + 1, nx - 1]
    for a in (gx, gy):
        a[0, 0] = 2.0 * a[1, 1] - a[2, 2]
        a[ny + 1, 0] = 2.0 * a[ny, 1] - a[ny - 1, 2]
        a[0, nx + 1] = 2.0 * a[1, nx] - a[2, nx - 1]
        a[ny + 1, nx + 1] = 2.0 * a[ny, nx] - a[ny - 1, nx - 1]

    # grid.f:FullGrid 312-362
    xu = field_array(ld, ncol); yu = field_array(ld, ncol)
    xv = field_array(ld, ncol); yv = field_array(ld, ncol)
    xc = field_array(ld, ncol); yc = field_array(ld, ncol)
    J1 = slice(1, ny + 2); J0 = slice(0, ny + 1)
    I = slice(0, nx + 2)
    xu[J1, I] = 0.5 * (gx[J1, I] + gx[J0, I])
    yu[J1, I] = 0.5 * (gy[J1, I] + gy[J0, I])
    Ja = slice(0, ny + 2); I1 = slice(1, nx + 2); I0 = slice(0, nx + 1)
    xv[Ja, I1] = 0.5 * (gx[Ja, I1] + gx[Ja, I0])
    yv[Ja, I1] = 0.5 * (gy[Ja, I1] + gy[Ja, I0])
    xc[J1, I1] = 0.5 * (xu[J1, I1] + xu[J1, I0])
    yc[J1, I1] = 0.5 * (yv[J1, I1] + yv[J0, I1])

    # grid.f:Metric 368-535 on i=1..nx, j=1..ny
    m = {nm: field_array(ld, ncol) for nm in METRIC_NAMES}
    J = slice(1, ny + 1); Jp = slice(2, ny + 2); Jm = slice(0, ny)
    Ic = slice(1, nx + 1); Ip = slice(2, nx + 2); Im = slice(0, nx)

    def inner(a, b, c, d):
        return a * b + c * d

    # N
    xzn = xv[J, Ip] - xv[J, Ic]; xen = xu[Jp, Ic] - xu[J, Ic]
    yzn = yv[J, Ip] - yv[J, Ic]; yen = yu[Jp, Ic] - yu[J, Ic]
    djn = 1.0 / (xzn * yen - xen * yzn)
    g11 = inner(xzn, xzn, yzn, yzn); g12 = inner(xzn, xen, yzn, yen); g22 = inner(xen, xen, yen, yen)
    m["xzn"][J, Ic], m["xen"][J, Ic], m["yzn"][J, Ic], m["yen"][J, Ic] = xzn, xen, yzn, yen
    m["djn"][J, Ic] = djn
    m["ran"][J, Ic] = djn * g22
    m["rbn"][J, Ic] = -djn * g12 * 0.25
    m["rgn"][J, Ic] = djn * g11
    # U
    xzu = xc[J, Ip] - xc[J, Ic]; xeu = gx[J, Ic] - gx[Jm, Ic]
    yzu = yc[J, Ip] - yc[J, Ic]; yeu = gy[J, Ic] - gy[Jm, Ic]
    dju = 1.0 / (xzu * yeu - xeu * yzu)
    g12 = inner(xzu, xeu, yzu, yeu); g22 = inner(xeu, xeu, yeu, yeu)
    m["xzu"][J, Ic], m["xeu"][J, Ic], m["yzu"][J, Ic], m["yeu"][J, Ic] = xzu, xeu, yzu, yeu
    m["dju"][J, Ic] = dju
    m["rau"][J, Ic] = dju * g22
    m["rbu"][J, Ic] = -dju * g12 * 0.25
    # V
    xzv = gx[J, Ic] - gx[J, Im]; xev = xc[Jp, Ic] - xc[J, Ic]
    yzv = gy[J, Ic] - gy[J, Im]; yev = yc[Jp, Ic] - yc[J, Ic]
    djv = 1.0 / (xzv * yev - xev * yzv)
    g11 = inner(xzv, xzv, yzv, yzv); g12 = inner(xzv, xev, yzv, yev)
    m["xzv"][J, Ic], m["xev"][J, Ic], m["yzv"][J, Ic], m["yev"][J, Ic] = xzv, xev, yzv, yev
    m["djv"][J, Ic] = djv
    m["rbv"][J, Ic] = -djv * g12 * 0.25
    m["rgv"][J, Ic] = djv * g11
    # C
    xzc = xu[J, Ic] - xu[J, Im]; xec = xv[J, Ic] - xv[Jm, Ic]
    yzc = yu[J, Ic] - yu[J, Im]; yec = yv[J, Ic] - yv[Jm, Ic]
    djc = 1.0 / (xzc * yec - xec * yzc)
    g11 = inner(xzc, xzc, yzc, yzc); g12 = inner(xzc, xec, yzc, yec); g22 = inner(xec, xec, yec, yec)
    m["xzc"][J, Ic], m["xec"][J, Ic], m["yzc"][J, Ic], m["yec"][J, Ic] = xzc, xec, yzc, yec
    m["djc"][J, Ic] = djc
    m["rac"][J, Ic] = djc * g22
    m["rbc"][J, Ic] = -djc * g12 * 0.25
    m["rgc"][J, Ic] = djc * g11
    return Grid(nx=nx, ny=ny, ld=ld, ncol=ncol, metrics=m, gx=gx, gy=gy)


def uniform_nodes(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0):
    """Nodes x=(i-1)/(nx-1)*Lx, y=(j-1)/(ny-1)*Ly as (ny, nx) arrays."""
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    x = np.broadcast_to(i[None, :] / (nx - 1) * lx, (ny, nx)).copy()
    y = np.broadcast_to(j[:, None] / (ny - 1) * ly, (ny, nx)).copy()
    return x, y


def skewed_nodes(nx: int, ny: int, stretch: float = 0.15, skew: float = 0.2, seed: int = 0):
    """A stretched, sheared, slightly perturbed grid: every cross metric is non-zero."""
    rng = np.random.default_rng(seed)
    s = np.arange(nx, dtype=np.float64) / (nx - 1)
    t = np.arange(ny, dtype=np.float64) / (ny - 1)
    xs = s + stretch * np.sin(np.pi * s) * s
    yt = t + 0.5 * stretch * np.sin(np.pi * t)
    x = xs[None, :] + skew * yt[:, None] * (1.0 + 0.1 * s[None, :])
    y = yt[:, None] * (1.0 + 0.05 * np.cos(np.pi * s[None, :])) + 0.03 * s[None, :]
    jit = 0.02 / max(nx, ny)
    x = x + jit * rng.standard_normal((ny, nx)) * (np.arange(ny)[:, None] % (ny - 1) > 0) * (np.arange(nx)[None, :] % (nx - 1) > 0)
    y = y + jit * rng.standard_normal((ny, nx)) * (np.arange(ny)[:, None] % (ny - 1) > 0) * (np.arange(nx)[None, :] % (nx - 1) > 0)
    return x, y


@dataclass
class Regions:
    """Region and boundary tables as built by bound_cond.f:SetUpBCs 34-443.

    All tables are Fortran-ordered int32/float64 arrays of the reference's
    declared extents (mgri, mgrj): nRegBrd(mgri,mgrj,4), nRegType(mgri,mgrj),
    nMomBdTp(mgri,mgrj,4), dBCVal(mgri,mgrj,4,4), dPR*(mgri,mgrj).

    ``profiles`` is an extension with no counterpart in the reference, which
    holds one constant per region side and variable: {(ir, jr, side, var):
    values} replaces the constant dBCVal(ir,jr,side,var) of u, v or p by one
    value per grid line along the side, indexed by j = 0..ny+1 on the west and
    east sides and by i = 0..nx+1 on the south and north sides.
    """
    nx: int
    ny: int
    nreg: np.ndarray
    nRegBrd: np.ndarray
    nRegType: np.ndarray
    nMomBdTp: np.ndarray
    dBCVal: np.ndarray
    dPRporos: np.ndarray
    dPRporc1: np.ndarray
    dPRporc2: np.ndarray
    mgri: int = MGRI
    mgrj: int = MGRJ
    profiles: dict = field(default_factory=dict)

    def border(self, ir, jr, side):
        return int(self.nRegBrd[ir - 1, jr - 1, side - 1])

    def btype(self, ir, jr, side):
        return int(self.nMomBdTp[ir - 1, jr - 1, side - 1])

    def rtype(self, ir, jr):
        return int(self.nRegType[ir - 1, jr - 1])

    def profile_length(self, side: int) -> int:
        return self.ny + 2 if side in (WEST, EAST) else self.nx + 2

    def set_profile(self, ir: int, jr: int, side: int, var: int, values) -> None:
        """Give side ``side`` of region (ir, jr) a profile of variable ``var``
        (VAR_U, VAR_V or VAR_P); ``values`` None removes it."""
        if not (1 <= ir <= self.nreg[0] and 1 <= jr <= self.nreg[1]):
            raise ValueError(f"region {ir},{jr} outside the {self.nreg[0]}x{self.nreg[1]} regions")
        if side not in (WEST, EAST, SOUTH, NORTH) or var not in (VAR_U, VAR_V, VAR_P):
            raise ValueError(f"bad side {side} or variable {var}")
        key = (ir, jr, side, var)
        if values is None:
            self.profiles.pop(key, None)
            return
        a = np.ascontiguousarray(values, dtype=np.float64)
        if a.shape != (self.profile_length(side),):
            raise ValueError(f"profile of side {side} needs {self.profile_length(side)} values, got {a.shape}")
        self.profiles[key] = a.copy()

    def profile_tables(self):
        """(poff, pval): poff(mgri,mgrj,4,3) Fortran-ordered int32, -1 where a
        side uses its constant, else the offset of its values in pval; both
        None without profiles."""
        if not self.profiles:
            return None, None
        poff = np.full((self.mgri, self.mgrj, 4, 3), -1, dtype=np.int32, order="F")
        parts, off = [], 0
        for (ir, jr, side, var), a in sorted(self.profiles.items()):
            poff[ir - 1, jr - 1, side - 1, var - 1] = off
            parts.append(a)
            off += a.size
        return poff, np.concatenate(parts)


def side_positions(x: np.ndarray, y: np.ndarray, reg: Regions, ir: int, jr: int, side: int,
                   var: int) -> np.ndarray:
    """Where each grid line's boundary value of a region side applies, as the
    fraction s of the side's length (0 at its first border node, 1 at its
    last), for k = 0..ny+1 (west/east sides) or 0..nx+1 (south/north).

    Along a west or east side u and p sit at cell centres (between nodes j-1
    and j) and v on node j; along a south or north side v and p sit at cell
    centres (between nodes i-1 and i) and u on node i.  Lengths are measured
    along the side's grid line; the ghost lines 0 and n+1 lie one spacing
    beyond the end nodes.  x, y are the (ny, nx) node arrays."""
    if side in (WEST, EAST):
        col = reg.border(ir, jr, side) - 1
        px, py = x[:, col], y[:, col]
        lo, hi = reg.border(ir, jr, SOUTH), reg.border(ir, jr, NORTH)
        centred = var in (VAR_U, VAR_P)
    else:
        row = reg.border(ir, jr, side) - 1
        px, py = x[row, :], y[row, :]
        lo, hi = reg.border(ir, jr, WEST), reg.border(ir, jr, EAST)
        centred = var in (VAR_V, VAR_P)
    arc = np.concatenate(([0.0], np.cumsum(np.hypot(np.diff(px), np.diff(py)))))
    n = arc.size                                  # nodes 1..n at arc[0..n-1]
    node = np.empty(n + 2)                        # node index 0..n+1
    node[1:n + 1] = arc
    node[0] = 2.0 * arc[0] - arc[1]
    node[n + 1] = 2.0 * arc[-1] - arc[-2]
    if centred:                                   # line k between nodes k-1 and k
        pos = np.empty(n + 2)
        pos[1:] = 0.5 * (node[:-1] + node[1:])
        pos[0] = 2.0 * node[0] - pos[1]
    else:
        pos = node
    return (pos - node[lo]) / (node[hi] - node[lo])


def parabolic_profile(s: np.ndarray, peak: float) -> np.ndarray:
    """4 peak s (1 - s): zero at both ends of the side, ``peak`` at its middle."""
    return 4.0 * peak * s * (1.0 - s)


def sine_profile(s: np.ndarray, amplitudes) -> np.ndarray:
    """sum_k a_k sin(k pi s), k = 1..len(amplitudes)."""
    out = np.zeros_like(s)
    for k, a in enumerate(amplitudes, start=1):
        out += a * np.sin(k * np.pi * s)
    return out


# wolfd2.h:24-31
RT_NOSRCE, RT_HEATGN, RT_TEMPER = 0, 1, 2
BT_INTERN, BT_TEMPER, BT_HTFLUX = 0, 1, 2


@dataclass
class Thermal:
    """Thermal region tables (parse.f:lReadBCFile, bound_cond.f:HeatGenST):
    nTRgType(mgri,mgrj) region types RT_*, nTemBdTp(mgri,mgrj,4) temperature
    boundary types BT_*, dTRgVal(mgri,mgrj) region temperatures, dHGSTval
    (mgri,mgrj) non-dimensional heat generation.  Boundary temperatures are
    the regions' dBCVal(.,.,side,VAR_T)."""
    nTRgType: np.ndarray
    nTemBdTp: np.ndarray
    dTRgVal: np.ndarray
    dHGSTval: np.ndarray


def make_thermal(reg: Regions, heat=(), fixed=(), walls=()) -> Thermal:
    """Thermal tables over ``reg`` (whose dBCVal receives the boundary
    temperatures): every region RT_NOSRCE with BT_INTERN sides and adiabatic
    (BT_HTFLUX, zero flux) outer walls (parse.f:InitBCFlags 2350-2375), then
    ``heat`` items (ireg, jreg, q) with q the non-dimensional source,
    ``fixed`` items (ireg, jreg, T) for fixed-temperature regions (their sides
    and the facing sides of neighbours become BT_TEMPER at T, as a 'blockage
    ... temper' line does, parse.f:1357-1392), and ``walls`` items (ireg,
    jreg, side, 'temper'|'flux', value)."""
    mgri, mgrj = reg.mgri, reg.mgrj
    ni, nj = int(reg.nreg[0]), int(reg.nreg[1])
    f = dict(order="F")
    rt = np.zeros((mgri, mgrj), dtype=np.int32, **f)
    bt = np.zeros((mgri, mgrj, 4), dtype=np.int32, **f)
    tv = np.zeros((mgri, mgrj), dtype=np.float64, **f)
    hg = np.zeros((mgri, mgrj), dtype=np.float64, **f)
    bt[0, :nj, WEST - 1] = BT_HTFLUX
    bt[ni - 1, :nj, EAST - 1] = BT_HTFLUX
    bt[:ni, 0, SOUTH - 1] = BT_HTFLUX
    bt[:ni, nj - 1, NORTH - 1] = BT_HTFLUX
    for ir, jr, q in heat:
        rt[ir - 1, jr - 1] = RT_HEATGN
        hg[ir - 1, jr - 1] = q
    for ir, jr, temp in fixed:
        rt[ir - 1, jr - 1] = RT_TEMPER
        tv[ir - 1, jr - 1] = temp
        for side in (WEST, EAST, SOUTH, NORTH):
            bt[ir - 1, jr - 1, side - 1] = BT_TEMPER
            reg.dBCVal[ir - 1, jr - 1, side - 1, VAR_T - 1] = temp
        for di, dj, side in ((-1, 0, EAST), (1, 0, WEST), (0, -1, NORTH), (0, 1, SOUTH)):
            a, b = ir + di, jr + dj
            if 1 <= a <= ni and 1 <= b <= nj:
                bt[a - 1, b - 1, side - 1] = BT_TEMPER
                reg.dBCVal[a - 1, b - 1, side - 1, VAR_T - 1] = temp
    for ir, jr, side, kind, value in walls:
        sd = SIDES[side] if isinstance(side, str) else side
        bt[ir - 1, jr - 1, sd - 1] = BT_TEMPER if kind == "temper" else BT_HTFLUX
        reg.dBCVal[ir - 1, jr - 1, sd - 1, VAR_T - 1] = value
    return Thermal(nTRgType=rt, nTemBdTp=bt, dTRgVal=tv, dHGSTval=hg)


def make_regions(nx: int, ny: int, nreg=(1, 1), i_borders=(), j_borders=(),
                 walls=(), inlets=(), outlets=(), blockages=(), porous=(),
                 mgri: int = MGRI, mgrj: int = MGRJ) -> Regions:
    """Region tables for ``nreg`` regions, as the boundary_conditions section builds them.

    ``walls`` items: (ireg, jreg, side, kind, tangent_vel, press) with kind
    'no_slip'|'no_stress' (parse.f TWALL); ``inlets``: (ireg, jreg, side, u, v,
    press); ``outlets``: (ireg, jreg, side, kind 'fully_dev'|'mass_cons', u, v,
    press); ``blockages``: (ireg, jreg); ``porous``: (ireg, jreg, poros, c1, c2)
    with already non-dimensional coefficients.  Velocities are non-dimensional
    (already divided by uref, parse.f:1532).
    """
    ni, nj = nreg
    if not (1 <= ni <= mgri and 1 <= nj <= mgrj):
        raise ValueError("too many regions")
    f = dict(order="F")
    brd = np.zeros((mgri, mgrj, 4), dtype=np.int32, **f)
    rtype = np.zeros((mgri, mgrj), dtype=np.int32, **f)
    btp = np.zeros((mgri, mgrj, 4), dtype=np.int32, **f)
    bcv = np.zeros((mgri, mgrj, 4, 4), dtype=np.float64, **f)
    poros = np.zeros((mgri, mgrj), dtype=np.float64, **f)
    porc1 = np.zeros((mgri, mgrj), dtype=np.float64, **f)
    porc2 = np.zeros((mgri, mgrj), dtype=np.float64, **f)

    for k, b in enumerate(i_borders):
        brd[k + 1, 0, WEST - 1] = b
    for k, b in enumerate(j_borders):
        brd[0, k + 1, SOUTH - 1] = b
    # parse.f:InitBCFlags 2257-2379
    rtype[:ni, :nj] = RM_INTERN
    btp[:ni, :nj, :] = BM_INTERN
    btp[0, :nj, WEST - 1] = BM_WALL1
    btp[ni - 1, :nj, EAST - 1] = BM_WALL1
    btp[:ni, 0, SOUTH - 1] = BM_WALL1
    btp[:ni, nj - 1, NORTH - 1] = BM_WALL1

    for (ir, jr) in blockages:        # parse.f TBLOCKAGE
        rtype[ir - 1, jr - 1] = RM_BLOCKG
        btp[ir - 1, jr - 1, :] = BM_WALL1
        if ir > 1:
            btp[ir - 2, jr - 1, EAST - 1] = BM_WALL1
        if ir < ni:
            btp[ir, jr - 1, WEST - 1] = BM_WALL1
        if jr > 1:
            btp[ir - 1, jr - 2, NORTH - 1] = BM_WALL1
        if jr < nj:
            btp[ir - 1, jr, SOUTH - 1] = BM_WALL1
    for (ir, jr, por, c1, c2) in porous:
        rtype[ir - 1, jr - 1] = RM_POROUS
        poros[ir - 1, jr - 1] = por
        porc1[ir - 1, jr - 1] = c1
        porc2[ir - 1, jr - 1] = c2
    for (ir, jr, side, kind, tv, pr) in walls:
        s = SIDES[side] if isinstance(side, str) else side
        btp[ir - 1, jr - 1, s - 1] = BM_WALL1 if kind == "no_slip" else BM_WALL2
        if s in (WEST, EAST):
            bcv[ir - 1, jr - 1, s - 1, VAR_V - 1] = tv
        else:
            bcv[ir - 1, jr - 1, s - 1, VAR_U - 1] = tv
        bcv[ir - 1, jr - 1, s - 1, VAR_P - 1] = pr
    for (ir, jr, side, u, v, pr) in inlets:
        s = SIDES[side] if isinstance(side, str) else side
        btp[ir - 1, jr - 1, s - 1] = BM_INLET
        bcv[ir - 1, jr - 1, s - 1, VAR_U - 1] = u
        bcv[ir - 1, jr - 1, s - 1, VAR_V - 1] = v
        bcv[ir - 1, jr - 1, s - 1, VAR_P - 1] = pr
    for (ir, jr, side, kind, u, v, pr) in outlets:
        s = SIDES[side] if isinstance(side, str) else side
        btp[ir - 1, jr - 1, s - 1] = BM_OUTLT1 if kind == "fully_dev" else BM_OUTLT2
        bcv[ir - 1, jr - 1, s - 1, VAR_U - 1] = u
        bcv[ir - 1, jr - 1, s - 1, VAR_V - 1] = v
        bcv[ir - 1, jr - 1, s - 1, VAR_P - 1] = pr

    # bound_cond.f:165-193 border assignment
    for jr in range(1, nj + 1):
        brd[0, jr - 1, WEST - 1] = 1
        brd[ni - 1, jr - 1, EAST - 1] = nx
    for ir in range(1, ni + 1):
        brd[ir - 1, 0, SOUTH - 1] = 1
        brd[ir - 1, nj - 1, NORTH - 1] = ny
    for jr in range(2, nj + 1):
        for ir in range(2, ni + 1):
            brd[ir - 1, jr - 1, WEST - 1] = brd[ir - 1, 0, WEST - 1]
            brd[ir - 1, jr - 1, SOUTH - 1] = brd[0, jr - 1, SOUTH - 1]
    for jr in range(1, nj + 1):
        for ir in range(1, ni):
            brd[ir - 1, jr - 1, EAST - 1] = brd[ir, 0, WEST - 1]
    for jr in range(1, nj):
        for ir in range(1, ni + 1):
            brd[ir - 1, jr - 1, NORTH - 1] = brd[0, jr, SOUTH - 1]
    # bound_cond.f:196-230 consistency of inlet/wall flags across interfaces
    for jr in range(1, nj + 1):
        for ir in range(1, ni):
            if btp[ir - 1, jr - 1, EAST - 1] == BM_INLET:
                btp[ir, jr - 1, WEST - 1] = BM_INLET
            if btp[ir, jr - 1, WEST - 1] == BM_INLET:
                btp[ir - 1, jr - 1, EAST - 1] = BM_INLET
            if btp[ir - 1, jr - 1, EAST - 1] == BM_WALL1:
                btp[ir, jr - 1, WEST - 1] = BM_WALL1
            if btp[ir, jr - 1, WEST - 1] == BM_WALL1:
                btp[ir - 1, jr - 1, EAST - 1] = BM_WALL1
    for jr in range(1, nj):
        for ir in range(1, ni + 1):
            if btp[ir - 1, jr, SOUTH - 1] == BM_INLET:
                btp[ir - 1, jr - 1, NORTH - 1] = BM_INLET
            if btp[ir - 1, jr - 1, NORTH - 1] == BM_INLET:
                btp[ir - 1, jr, SOUTH - 1] = BM_INLET
            if btp[ir - 1, jr, SOUTH - 1] == BM_WALL1:
                btp[ir - 1, jr - 1, NORTH - 1] = BM_WALL1
            if btp[ir - 1, jr - 1, NORTH - 1] == BM_WALL1:
                btp[ir - 1, jr, SOUTH - 1] = BM_WALL1
    nreg_arr = np.array([ni, nj], dtype=np.int32)
    return Regions(nx=nx, ny=ny, nreg=nreg_arr, nRegBrd=brd, nRegType=rtype, nMomBdTp=btp,
                   dBCVal=bcv, dPRporos=poros, dPRporc1=porc1, dPRporc2=porc2, mgri=mgri, mgrj=mgrj)


def cavity_regions(nx: int, ny: int, lid: float = 1.0) -> Regions:
    """One region, no-slip walls, north wall moving at ``lid`` (lid-driven cavity)."""
    return make_regions(nx, ny, walls=[(1, 1, "north", "no_slip", lid, 0.0)])


@dataclass
class Params:
    """Non-dimensional run parameters (file_manip.f:ReadParam 120-159, parse.f:117-193)."""
    dk: float
    re: float
    fr: float = 1.0 / 9.81
    mqiter: int = 20
    qtol: float = 1e-4
    nPpeSolver: int = 1
    msorit: int = 2000
    sortol: float = 1e-8
    sorrel: float = 1.0
    cartesian: bool = True

    @staticmethod
    def from_physical(dt: float, rho: float, mu: float, dlref: float = 1.0, uref: float = 1.0,
                      **kw) -> "Params":
        dk = dt * uref / dlref                     # file_manip.f:134
        dkinvisc = mu / rho                        # file_manip.f:145
        re = dlref * uref / dkinvisc               # file_manip.f:157
        fr = uref ** 2.0 / (dlref * 9.81)          # file_manip.f:159
        return Params(dk=dk, re=re, fr=fr, **kw)


@dataclass
class StepOptions:
    """The rest of a time step's switches (parse.f:117-193 defaults): the
    energy equation, the equation of state (buoyancy), the momentum/energy
    iterations and the Shuman filters, with the physical constants they use."""
    nthermen: int = 0
    neqstate: int = 0
    nmeiter: int = 1
    nfiltu: int = 0
    nfiltv: int = 0
    nfiltt: int = 0
    dmeittol: float = 1.0e-4
    fpu: float = 5.0e2
    fpv: float = 5.0e2
    fpt: float = 5.0
    pe: float = 1.0
    uref: float = 1.0
    densref: float = 1.0
    tmax: float = 1.0e3
    tref: float = 2.8357e2
    rconst: float = 0.0


# the ppe_solver keywords (parse.f:446-466); "mg" (7) is the multigrid solver of this path,
# which the reference does not have
PPE_SOLVERS = {"sor": 1, "lsor": 2, "rb_lsor": 3, "par_rb_lsor": 4, "rb_sor": 5, "par_rb_sor": 6, "mg": 7}
