"""Input decks, grids and restart files in the reference solver's formats.  TEST INFRASTRUCTURE
(drives the compiled reference as a checker and CPU baseline).

Writes what the reference program reads (the keyword deck of
src/parse.f:lReadIPFile 32-878 and lReadBCFile 1004-1951, the fluid table of
src/parse.f:lReadFldProps 2027-2249, the 2-D Plot3D grid of
src/grid.f:ReadGrid 150-252) and reads back what it writes (the unformatted
restart file of src/file_manip.f:SaveRestart 364-401 and the per-step table of
src/string.f:PrintDiff 467-570).  Test infrastructure only.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

import numpy as np


def write_grid(path: str, x: np.ndarray, y: np.ndarray) -> None:
    """Plot3D 2-D grid: 'nx ny' then x(i,j) i fastest, then y (grid.f:189-228).

    x, y have shape (ny, nx) (row j, column i).  Values are written with 17
    significant digits so the list-directed read gives back the same doubles.
    """
    ny, nx = x.shape
    with open(path, "w") as f:
        f.write(f"{nx} {ny}\n")
        for arr in (x, y):
            flat = arr.reshape(-1)
            for k in range(0, flat.size, 4):
                f.write(" ".join(f"{v:.17e}" for v in flat[k:k + 4]) + "\n")


def uniform_grid(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0):
    """x=(i-1)/(nx-1)*Lx, y=(j-1)/(ny-1)*Ly on nx*ny nodes (shape (ny, nx))."""
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    x = np.broadcast_to(i[None, :] / (nx - 1) * lx, (ny, nx)).copy()
    y = np.broadcast_to(j[:, None] / (ny - 1) * ly, (ny, nx)).copy()
    return x, y


def write_fluid_table(path: str, name: str, rho: float, mu: float, gas_constant: float | None = None) -> None:
    """Two identical rows bracketing tref, so the interpolation is exact; the
    gas constant (parse.f:2134-2143) is given on the 'fluid' line."""
    with open(path, "w") as f:
        f.write(f"fluid {name}" + (f" gasconstant {gas_constant!r}" if gas_constant is not None else "") + "\n")
        for t in (200.0, 400.0):
            f.write(f"{t:.1f} {rho!r} 1.0 {mu!r} 0.025 0.7\n")
        f.write("end\n")


@dataclass
class Deck:
    """One run of the reference program (keywords of parse.f:905-973, 1978-1996)."""
    nts: int
    dt: float
    mu: float
    rho: float = 1.0
    grid_file: str = "grid.xyz"
    cartesian: bool = True
    ppe_solver: str = "sor"
    max_ql_iter: int | None = None
    ql_tolerance: float | None = None
    max_sor_iter: int | None = None
    sor_tolerance: float | None = None
    sor_relaxation: float | None = None
    restart_in: str | None = None
    restart_out: str = "out.rst"
    print_diff_freq: int = 1
    nreg: tuple = (1, 1)
    i_borders: tuple = ()
    j_borders: tuple = ()
    bc_lines: list = field(default_factory=lambda: ["wall 1 1 north tangent_vel 1.0"])
    extra_input: list = field(default_factory=list)   # more input_parameters lines
    gas_constant: float | None = None

    def input_text(self) -> str:
        grid = f"grid_file {self.grid_file}" + (" cartesian_grid" if self.cartesian else "")
        kw = [
            "section input_parameters",
            grid,
            f"n_time_steps {self.nts}",
            f"time_step_size {self.dt!r}",
            "fluid_prop_file fluid.dat",
            "fluid_prop_table fl01",
            f"ppe_solver {self.ppe_solver}",
        ]
        if self.max_ql_iter is not None:
            kw.append(f"max_ql_iter {self.max_ql_iter}")
        if self.ql_tolerance is not None:
            kw.append(f"ql_tolerance {self.ql_tolerance!r}")
        if self.max_sor_iter is not None:
            kw.append(f"max_sor_iter {self.max_sor_iter}")
        if self.sor_tolerance is not None:
            kw.append(f"sor_tolerance {self.sor_tolerance!r}")
        if self.sor_relaxation is not None:
            kw.append(f"sor_relaxation {self.sor_relaxation!r}")
        kw.extend(self.extra_input)
        if self.restart_in:
            kw.append(f"restart {self.restart_in}")
        kw.append(f"write_restart {self.restart_out}")
        kw.append("output_prefix out")
        kw.append(f"print_diff_freq {self.print_diff_freq}")
        # main.f:855 evaluates mod(k-ks, isfreq) even when no snapshots are
        # requested (amdflang does not short-circuit .and.); isfreq has no default.
        kw.append("snapshots_freq 1000000")
        kw.append("end section")
        kw.append("")
        kw.append("section boundary_conditions")
        kw.append(f"number_of_regions {self.nreg[0]} {self.nreg[1]}")
        kw.append("i_borders " + " ".join(str(b) for b in self.i_borders))
        kw.append("j_borders " + " ".join(str(b) for b in self.j_borders))
        kw.extend(self.bc_lines)
        kw.append("end section")
        return "\n".join(kw) + "\n"

    def write(self, rundir: str, x: np.ndarray, y: np.ndarray) -> None:
        os.makedirs(rundir, exist_ok=True)
        write_grid(os.path.join(rundir, self.grid_file), x, y)
        write_fluid_table(os.path.join(rundir, "fluid.dat"), "fl01", self.rho, self.mu, self.gas_constant)
        with open(os.path.join(rundir, "input.dat"), "w") as f:
            f.write(self.input_text())


# the restart format is the product's own reader/writer (wolfd2_amd.deck)
from wolfd2_amd.deck import read_restart, write_restart  # noqa: E402,F401


_DIFF_RE = re.compile(r"^\s*(\d+)\s*([-+0-9.Ee]+)\s*(-?\d+)([ *])\s*(\d+)\s+([-+0-9.Ee]+)\s+([-+0-9.Ee]+)\s+([-+0-9.Ee]+)")


def parse_print_diff(text: str) -> np.ndarray:
    """Rows (step, time, ql_iter, ppe_iter, |p1-pn|, |u1-un|, |v1-vn|) of the PrintDiff table."""
    rows = []
    for line in text.splitlines():
        m = _DIFF_RE.match(line)
        if m:
            rows.append([float(m.group(1)), float(m.group(2)), float(m.group(3)),
                         float(m.group(5)), float(m.group(6)), float(m.group(7)),
                         float(m.group(8))])
    return np.array(rows, dtype=np.float64).reshape(-1, 7)
