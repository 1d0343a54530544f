"""The workloads the path is measured and checked on, as (grid, regions, params).

* ``cavity``: the lid-driven square cavity (one region, no-slip walls, the
  north wall moving), fluid at rest, cold start.
* ``channel``: a channel of aspect Lx:1 with a parabolic inflow of peak 1 on
  the west side (a boundary profile, bound_cond.f:585-593 holds only a
  constant inflow per region side), no-slip south and north walls and an
  outlet on the east side: ``fully_dev`` (zero gradient, bound_cond.f:1656-1874
  OUTLT1) or ``mass_cons`` (OUTLT2).
* ``poiseuille_noise``: the channel's initial state, u = 4y(1-y) (the inflow
  profile) plus seeded Gaussian noise on the interior u and v, given to a run
  through a restart file of the reference's format (file_manip.f:409-463), the
  only way the reference takes a non-zero initial field.

Physical inputs are scaled as file_manip.f:ReadParam does (rho = 1,
uref = dlref = 1, so Re = 1/mu and dk = dt).
"""
from __future__ import annotations

import os
import tempfile

import numpy as np

from . import case, deck


def cavity(n: int, re: float, dt: float, lid: float = 1.0, **prm_kw):
    """n x n cells (n+1 nodes per side) on the unit square."""
    nx = ny = n + 1
    x, y = case.uniform_nodes(nx, ny)
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(nx, ny, lid=lid)
    prm = case.Params.from_physical(dt, 1.0, 1.0 / re, **prm_kw)
    return g, reg, prm


def channel(ncx: int, ncy: int, lx: float = 4.0, re: float = 500.0, dt: float = 1e-4, outlet: str = "fully_dev",
            peak: float = 1.0, **prm_kw):
    """ncx x ncy cells on [0, lx] x [0, 1]; returns (g, reg, prm, x, y)."""
    nx, ny = ncx + 1, ncy + 1
    x, y = case.uniform_nodes(nx, ny, lx=lx)
    g = case.compute_grid(x, y)
    reg = case.make_regions(nx, ny, inlets=[(1, 1, "west", peak, 0.0, 0.0)],
                            outlets=[(1, 1, "east", outlet, 0.0, 0.0, 0.0)])
    s = case.side_positions(x, y, reg, 1, 1, case.WEST, case.VAR_U)
    reg.set_profile(1, 1, case.WEST, case.VAR_U, case.parabolic_profile(s, peak))
    prm = case.Params.from_physical(dt, 1.0, 1.0 / re, **prm_kw)
    return g, reg, prm, x, y


def poiseuille_noise(g, reg, x, y, amp: float = 1e-3, seed: int = 20261202, peak: float = 1.0):
    """u = 4 peak y(1-y) on every u line j = 1..ny+1 (y of the cell centre
    between nodes j-1 and j, which is the inflow profile's position rule; on
    the ghost lines j = 1 and ny+1 the parabola's continuation, -u of the
    line inside up to O(h^2)) for i = 1..nx+1; N(0, amp^2) noise added on the
    interior u (i = 2..nx-1, j = 2..ny) and v (i = 2..nx, j = 2..ny-1), with
    the v ghost columns i = 1 and nx+1 set to minus their neighbours (zero v
    on the inflow and outflow faces); p = 0.  Fields are (ncol, ld) [j, i]
    arrays."""
    nx, ny = g.nx, g.ny
    s = case.side_positions(x, y, reg, 1, 1, case.WEST, case.VAR_U)   # k = j = 0..ny+1
    prof = case.parabolic_profile(s, peak)
    u = np.zeros((g.ncol, g.ld))
    v = np.zeros((g.ncol, g.ld))
    p = np.zeros((g.ncol, g.ld))
    u[1:ny + 2, 1:nx + 2] = prof[1:ny + 2, None]
    rng = np.random.default_rng(seed)
    u[2:ny + 1, 2:nx] += amp * rng.standard_normal((ny - 1, nx - 2))
    v[2:ny, 2:nx + 1] += amp * rng.standard_normal((ny - 2, nx - 1))
    v[2:ny, 1] = -v[2:ny, 2]
    v[2:ny, nx + 1] = -v[2:ny, nx]
    return u, v, p


def through_restart(g, u, v, p, k: int = 0, dtime: float = 0.0):
    """Write u, v, p to a restart file of the reference's format and read them
    back (the file the reference's InitialCond would read), returning the
    fields as (ncol, ld) arrays."""
    nx, ny = g.nx, g.ny
    S = np.s_[: ny + 2, : nx + 2]
    with tempfile.TemporaryDirectory(prefix="wd2_rst_") as d:
        path = os.path.join(d, "init.rst")
        deck.write_restart(path, k, dtime, nx, ny, {"u": u[S], "v": v[S], "p": p[S]})
        r = deck.read_restart(path)
    out = []
    for name in ("u", "v", "p"):
        f = np.zeros((g.ncol, g.ld))
        f[S] = r[name]
        out.append(f)
    return tuple(out)


def ensemble_specs(n: int = 64, seed: int = 20261202, sides=tuple(range(128, 769, 64)),
                   re_range=(100.0, 2000.0)):
    """The ensemble of cavities: n (side, Re) pairs, sides drawn uniformly from
    {128, 192, ..., 768} cells and Re uniformly from [100, 2000], seeded."""
    rng = np.random.default_rng(seed)
    side = rng.choice(np.asarray(sides), size=n)
    re = rng.uniform(re_range[0], re_range[1], size=n)
    return [(int(a), float(b)) for a, b in zip(side, re)]


def lid_modes(x, y, reg, n_modes: int = 16, sigma: float = 0.1, seed: int = 20261203, base: float = 1.0):
    """A sinusoidal lid profile: base + sum_m a_m sin(m pi s), m = 1..n_modes,
    a_m ~ N(0, sigma^2) (seeded), s the fraction of the lid's length at each
    u line of the north side (case.side_positions); returned for i = 0..nx+1
    (the lid values are i = 2..nx)."""
    s = case.side_positions(x, y, reg, 1, 1, case.NORTH, case.VAR_U)
    a = np.random.default_rng(seed).normal(0.0, sigma, size=n_modes)
    m = np.arange(1, n_modes + 1)
    return base + np.sin(np.pi * np.outer(s, m)) @ a


def adjoint_cavity(n: int = 512, re: float = 400.0, dt: float = 2e-4, qtol: float = -1.0, mqiter: int = 24,
                   sortol: float = 1e-13, msorit: int = 100):
    """The adjoint case: an n x n cavity (Re 400), FP64, every q-l loop run
    for mqiter iterations (a negative qtol never stops a loop; the increments
    reach round-off in 24 iterations at this dt) and multigrid Poisson solves
    converged tightly, so the run is the converged map the adjoint
    differentiates; returns (g, reg, prm, x, y).

    No q-l tolerance is used because the reference's stopping norm
    (DiffMaxNorm, utility.f) skips the rows next to the north and south
    walls: in the first steps of an impulsively started lid the increments
    sit on row ny (from rest they are exactly zero elsewhere, the wall rows
    of the second split being identity rows), so the loop stops after one
    iteration with that row unconverged, even at qtol 0."""
    nx = ny = n + 1
    x, y = case.uniform_nodes(nx, ny)
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(nx, ny, lid=1.0)
    prm = case.Params.from_physical(dt, 1.0, 1.0 / re, nPpeSolver=7, msorit=msorit, sortol=sortol, qtol=qtol,
                                    mqiter=mqiter)
    return g, reg, prm, x, y


def interior_u(g):
    """The u points a step updates (i = 2..nx-1, j = 2..ny) as a [j, i] slice."""
    return np.s_[2:g.ny + 1, 2:g.nx]


def interior_v(g):
    """The v points a step updates (i = 2..nx, j = 2..ny-1) as a [j, i] slice."""
    return np.s_[2:g.ny, 2:g.nx + 1]
