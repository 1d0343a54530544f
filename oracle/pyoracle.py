"""ctypes bindings of the parity checkers.  TEST INFRASTRUCTURE ONLY.

Two checkers live under oracle/:

* ``Oracle`` -- liboracle.so, the C restatement of the hot path (wd2_oracle.c),
  built by ``make -C oracle oracle``.  Available wherever the repo is.
* ``Reference`` -- oracle/_ref/std/libwolfd2ref.so, the reference's own Fortran
  sources compiled by oracle/Makefile.  Its arrays are dimensioned by the
  compiled-in config.f (mnx = mny = 302, mgri = 20, mgrj = 10); it keeps
  gfortran-style static local arrays between calls, so each grid size gets its
  own private copy of the library (``Reference.for_grid``).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only to check or time against.  Fields are numpy float64 arrays of
shape (ncol, ld) indexed [j, i] (Fortran (0:ld-1, 0:ncol-1) column-major).
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_SO = os.path.join(REF_DIR, "std", "libwolfd2ref.so")
REF_MNX = 302
REF_MGRI, REF_MGRJ = 20, 10

METRIC_ORDER = (
    "ran", "rbn", "rgn", "rac", "rbc", "rgc", "rau", "rbu", "rbv", "rgv",
    "dju", "djv", "djc", "djn", "xen", "yen", "xzn", "yzn", "xec", "yec",
    "xzc", "yzc", "xeu", "yeu", "xzv", "yzv", "xzu", "yzu", "xev", "yev",
)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.c_contiguous or a.flags.f_contiguous)
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


class _Regions(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("ld", C.c_int), ("mgri", C.c_int), ("mgrj", C.c_int),
                ("nReg", _ip), ("nRegBrd", _ip), ("nRegType", _ip), ("nMomBdTp", _ip),
                ("dBCVal", _dp), ("dPRporos", _dp), ("dPRporc1", _dp), ("dPRporc2", _dp),
                ("poff", _ip), ("pval", _dp)]


class _Thermal(C.Structure):
    _fields_ = [("nTRgType", _ip), ("nTemBdTp", _ip), ("dTRgVal", _dp), ("dHGSTval", _dp)]


class _StepOptions(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("nthermen", "neqstate", "nmeiter", "nfiltu", "nfiltv", "nfiltt")] + \
        [(f, C.c_double) for f in ("dmeittol", "fpu", "fpv", "fpt", "pe", "uref", "densref", "tmax", "tref",
                                   "rconst")]


class _Metrics(C.Structure):
    _fields_ = [(nm, _dp) for nm in METRIC_ORDER]


def oracle_available() -> bool:
    return os.path.exists(ORACLE_SO)


def reference_available() -> bool:
    return os.path.exists(REF_SO)


class Oracle:
    """The C restatement (liboracle.so)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle oracle`")
        L = C.CDLL(path)
        self.L = L
        L.wd2o_dmax_norm.restype = C.c_double
        L.wd2o_diff_max_norm.restype = C.c_double
        self._keep = []

    # -- argument packing ----------------------------------------------------
    def regions(self, reg, ld: int):
        s = _Regions(reg.nx, reg.ny, ld, reg.mgri, reg.mgrj, _i(reg.nreg), _i(reg.nRegBrd),
                     _i(reg.nRegType), _i(reg.nMomBdTp), _d(reg.dBCVal), _d(reg.dPRporos),
                     _d(reg.dPRporc1), _d(reg.dPRporc2))
        poff, pval = reg.profile_tables()
        if poff is not None:
            s.keep = (poff, pval)   # alive as long as the structure
            s.poff, s.pval = _i(poff), _d(pval)
        return s

    def metrics(self, grid):
        return _Metrics(*[_d(grid.metrics[nm]) for nm in METRIC_ORDER])

    @staticmethod
    def thermal(th):
        return _Thermal(_i(th.nTRgType), _i(th.nTemBdTp), _d(th.dTRgVal), _d(th.dHGSTval))

    # -- thermal routines ----------------------------------------------------
    def temp_bound_cond(self, reg, th, ld, t):
        return self.L.wd2o_temp_bound_cond(C.byref(self.regions(reg, ld)), C.byref(self.thermal(th)), _d(t))

    def therm_energy(self, reg, th, grid, dk, pe, un, vn, u, v, tn, t):
        return self.L.wd2o_therm_energy(C.byref(self.regions(reg, grid.ld)), C.byref(self.thermal(th)),
                                        C.c_double(dk), C.c_double(pe), C.byref(self.metrics(grid)),
                                        _d(un), _d(vn), _d(u), _d(v), _d(tn), _d(t))

    def eq_state(self, grid, uref, densref, tmax, tref, rconst, p, t, den):
        self.L.wd2o_eq_state(grid.nx, grid.ny, grid.ld, *[C.c_double(x) for x in (uref, densref, tmax, tref, rconst)],
                             _d(p), _d(t), _d(den))

    def filter(self, reg, th, ld, ncomp, fp, q):
        return self.L.wd2o_filter(C.byref(self.regions(reg, ld)), C.byref(self.thermal(th)), ncomp,
                                  C.c_double(fp), _d(q))

    # -- hot-path routines ---------------------------------------------------
    def alt_trid_lu(self, a: np.ndarray, b: np.ndarray):
        """a: (n,3) rows (sub, diag, super); solved in place like AltTridLU."""
        a = np.array(a, dtype=np.float64, order='C', copy=True)
        b = np.array(b, dtype=np.float64, copy=True)
        self.L.wd2o_alt_trid_lu(C.c_int(b.size), _d(a), _d(b))
        return a, b

    def vel_bound_cond(self, reg, ld, u, v):
        return self.L.wd2o_vel_bound_cond(C.byref(self.regions(reg, ld)), _d(u), _d(v))

    def vel_outflow_bcs(self, reg, ld, u, v):
        return self.L.wd2o_vel_outflow_bcs(C.byref(self.regions(reg, ld)), _d(u), _d(v))

    def pres_bound_cond(self, reg, ld, p):
        return self.L.wd2o_pres_bound_cond(C.byref(self.regions(reg, ld)), _d(p))

    def project(self, reg, grid, dk, p, u, v):
        return self.L.wd2o_project(C.byref(self.regions(reg, grid.ld)), C.c_double(dk),
                                   C.byref(self.metrics(grid)), _d(p), _d(u), _d(v))

    def divergence(self, grid, nloc, u, v, div):
        return self.L.wd2o_divergence(grid.nx, grid.ny, grid.ld, nloc, _d(grid.xeu), _d(grid.yeu),
                                      _d(grid.xzv), _d(grid.yzv), _d(u), _d(v), _d(div))

    def ppe(self, reg, grid, prm, u, v, p, solver=None, msorit=None, sortol=None, sorrel=None):
        nconv = C.c_int(0)
        st = self.L.wd2o_ppe(C.byref(self.regions(reg, grid.ld)), C.c_int(int(prm.cartesian)),
                             C.c_int(prm.nPpeSolver if solver is None else solver),
                             C.c_int(prm.msorit if msorit is None else msorit), C.byref(nconv),
                             C.c_double(prm.dk), C.c_double(prm.sortol if sortol is None else sortol),
                             C.c_double(prm.sorrel if sorrel is None else sorrel),
                             C.byref(self.metrics(grid)), _d(u), _d(v), _d(p))
        assert st == 0
        return nconv.value

    def x_momentum(self, reg, grid, prm, us, vs, un, vn, dus):
        return self.L.wd2o_x_momentum(C.byref(self.regions(reg, grid.ld)), C.c_double(prm.dk),
                                      C.c_double(prm.re), C.byref(self.metrics(grid)),
                                      _d(us), _d(vs), _d(un), _d(vn), _d(dus))

    def y_momentum(self, reg, grid, prm, d, dn, us, vs, un, vn, dvs):
        return self.L.wd2o_y_momentum(C.byref(self.regions(reg, grid.ld)), C.c_double(prm.dk),
                                      C.c_double(prm.re), C.c_double(prm.fr), C.byref(self.metrics(grid)),
                                      _d(d), _d(dn), _d(us), _d(vs), _d(un), _d(vn), _d(dvs))

    def naux_momentum(self, reg, grid, prm, d, dn, un, vn, us, vs, mqiter=None, qtol=None):
        return self.L.wd2o_naux_momentum(C.byref(self.regions(reg, grid.ld)),
                                         C.c_int(prm.mqiter if mqiter is None else mqiter),
                                         C.c_double(prm.dk), C.c_double(prm.re), C.c_double(prm.fr),
                                         C.c_double(prm.qtol if qtol is None else qtol),
                                         C.byref(self.metrics(grid)), _d(d), _d(dn), _d(un), _d(vn),
                                         _d(us), _d(vs))

    def dmax_norm(self, grid, u):
        return self.L.wd2o_dmax_norm(grid.nx, grid.ny, grid.ld, _d(u))

    def diff_max_norm(self, grid, un, u):
        return self.L.wd2o_diff_max_norm(grid.nx, grid.ny, grid.ld, _d(un), _d(u))

    def cold_start(self, reg, grid, prm, u, v, p):
        n = C.c_int(0)
        st = self.L.wd2o_cold_start(C.byref(self.regions(reg, grid.ld)), C.byref(self.metrics(grid)),
                                    C.c_int(int(prm.cartesian)), C.c_int(prm.nPpeSolver),
                                    C.c_int(prm.msorit), C.c_double(prm.sortol), C.c_double(prm.sorrel),
                                    C.c_double(prm.dk), _d(u), _d(v), _d(p), C.byref(n))
        assert st == 0
        return n.value

    def run(self, reg, grid, prm, u, v, p, nsteps: int, trace: bool = False):
        """Advance nsteps of main.f:690-972 in place; returns per-step (ql, ppe, dp, du, dv)."""
        shape = u.shape
        work = [np.zeros(shape) for _ in range(5)]
        d = np.zeros(shape)
        dn = np.zeros(shape)
        it = (C.c_int * 2)()
        dif = (C.c_double * 3)()
        rs, ms = self.regions(reg, grid.ld), self.metrics(grid)
        rows = []
        for _ in range(nsteps):
            st = self.L.wd2o_step(C.byref(rs), C.byref(ms), C.c_int(int(prm.cartesian)),
                                  C.c_int(prm.mqiter), C.c_double(prm.qtol), C.c_int(prm.nPpeSolver),
                                  C.c_int(prm.msorit), C.c_double(prm.sortol), C.c_double(prm.sorrel),
                                  C.c_double(prm.dk), C.c_double(prm.re), C.c_double(prm.fr),
                                  _d(d), _d(dn), _d(u), _d(v), _d(p), *[_d(w) for w in work], it, dif)
            assert st == 0, st
            rows.append((it[0], it[1], dif[0], dif[1], dif[2]))
        return np.array(rows, dtype=np.float64)


def run_ex(oracle, reg, th, grid, prm, opt, u, v, p, t, d, nsteps: int):
    """Advance nsteps of main.f:690-972 with step options ``opt``
    (case.StepOptions) in place; returns rows (ql, ppe, dp, du, dv, dt)."""
    shape = u.shape
    work = [np.zeros(shape) for _ in range(7)]
    dn = np.zeros(shape)
    it = (C.c_int * 2)()
    dif = (C.c_double * 4)()
    rs, ms, ts = oracle.regions(reg, grid.ld), oracle.metrics(grid), oracle.thermal(th)
    o = _StepOptions(*[getattr(opt, f) for f, _ in _StepOptions._fields_])
    rows = []
    for _ in range(nsteps):
        st = oracle.L.wd2o_step_ex(C.byref(rs), C.byref(ts), C.byref(ms), C.byref(o), C.c_int(int(prm.cartesian)),
                                   C.c_int(prm.mqiter), C.c_double(prm.qtol), C.c_int(prm.nPpeSolver),
                                   C.c_int(prm.msorit), C.c_double(prm.sortol), C.c_double(prm.sorrel),
                                   C.c_double(prm.dk), C.c_double(prm.re), C.c_double(prm.fr),
                                   _d(d), _d(dn), _d(u), _d(v), _d(p), _d(t), *[_d(w) for w in work], it, dif)
        assert st == 0, st
        rows.append((it[0], it[1], dif[0], dif[1], dif[2], dif[3]))
    return np.array(rows, dtype=np.float64)


class Reference:
    """The reference's own compiled subroutines (libwolfd2ref.so, Fortran ABI).

    Arrays passed in must have the compiled-in shape (303, 303) ([j, i]).
    """
    LD = REF_MNX + 1
    _copies: dict = {}

    def __init__(self, path: str):
        self.L = C.CDLL(path)
        self.L.dmaxnorm_.restype = C.c_double
        self.L.diffmaxnorm_.restype = C.c_double

    @classmethod
    def for_grid(cls, key) -> "Reference":
        """A private copy of the library per grid (its static locals are per copy)."""
        if key not in cls._copies:
            if not os.path.exists(REF_SO):
                raise FileNotFoundError(f"{REF_SO} missing: run `make -C oracle ref`")
            tmp = tempfile.mkdtemp(prefix="wd2ref_")
            dst = os.path.join(tmp, f"libwolfd2ref_{len(cls._copies)}.so")
            shutil.copy(REF_SO, dst)
            cls._copies[key] = cls(dst)
        return cls._copies[key]

    @staticmethod
    def field():
        return np.zeros((REF_MNX + 1, REF_MNX + 1))

    @staticmethod
    def _ir(x):
        return C.byref(C.c_int(int(x)))

    @staticmethod
    def _dr(x):
        return C.byref(C.c_double(float(x)))

    def _reg(self, reg):
        assert reg.mgri == REF_MGRI and reg.mgrj == REF_MGRJ
        return [_i(reg.nreg), _i(reg.nRegBrd), _i(reg.nRegType), _i(reg.nMomBdTp)]

    def grid(self, nx, ny, gx, gy, out: dict):
        """MirrorPts + FullGrid + Metric (grid.f:258-535) on already scaled nodes."""
        L = self.L
        L.mirrorpts_(self._ir(nx), self._ir(ny), _d(gx), _d(gy))
        tmp = [self.field() for _ in range(6)]
        L.fullgrid_(self._ir(nx), self._ir(ny), _d(gx), _d(gy), *[_d(t) for t in tmp])
        names = ["rau", "rbu", "rbv", "rgv", "ran", "rbn", "rgn", "rac", "rbc", "rgc",
                 "dju", "djv", "djc", "djn", "xen", "yen", "xzn", "yzn", "xec", "yec", "xzc", "yzc",
                 "xeu", "yeu", "xzv", "yzv", "xzu", "yzu", "xev", "yev"]
        for nm in names:
            out.setdefault(nm, self.field())
        L.metric_(self._ir(nx), self._ir(ny), _d(gx), _d(gy), *[_d(t) for t in tmp],
                  *[_d(out[nm]) for nm in names])
        return out

    def alt_trid_lu(self, a: np.ndarray, b: np.ndarray):
        a = np.array(a, dtype=np.float64, order='C', copy=True)
        b = np.array(b, dtype=np.float64, copy=True)
        self.L.alttridlu_(self._ir(b.size), _d(a), _d(b))
        return a, b

    def vel_bound_cond(self, reg, u, v):
        r = self._reg(reg)
        self.L.velboundcond_(self._ir(reg.nx), self._ir(reg.ny), r[0], r[1], r[3], _d(reg.dBCVal), _d(u), _d(v))

    def vel_outflow_bcs(self, reg, u, v):
        r = self._reg(reg)
        self.L.veloutflowbcs_(self._ir(reg.nx), self._ir(reg.ny), r[0], r[1], r[3], _d(reg.dBCVal), _d(u), _d(v))

    def pres_bound_cond(self, reg, p):
        r = self._reg(reg)
        self.L.presboundcond_(self._ir(reg.nx), self._ir(reg.ny), *r, _d(reg.dBCVal), _d(p))

    def project(self, reg, g, dk, p, u, v):
        r = self._reg(reg)
        self.L.project_(self._ir(reg.nx), self._ir(reg.ny), *r, self._dr(dk), _d(g["dju"]), _d(g["djv"]),
                        _d(g["yeu"]), _d(g["xzv"]), _d(g["yzu"]), _d(g["xev"]), _d(p), _d(u), _d(v))

    def divergence(self, nx, ny, nloc, g, u, v, div):
        self.L.divergence_(self._ir(nx), self._ir(ny), self._ir(nloc), _d(g["xeu"]), _d(g["yeu"]),
                           _d(g["xzv"]), _d(g["yzv"]), _d(u), _d(v), _d(div))

    def ppe(self, reg, g, cartesian, solver, msorit, dk, sortol, sorrel, u, v, p):
        r = self._reg(reg)
        n = C.c_int(0)
        self.L.ppe_(self._ir(reg.nx), self._ir(reg.ny), r[0], r[1], r[2],
                    C.byref(C.c_int(1 if cartesian else 0)), self._ir(solver), self._ir(msorit), C.byref(n),
                    self._dr(dk), self._dr(sortol), self._dr(sorrel),
                    _d(g["rau"]), _d(g["rbu"]), _d(g["rbv"]), _d(g["rgv"]),
                    _d(g["xeu"]), _d(g["yeu"]), _d(g["xzv"]), _d(g["yzv"]), _d(u), _d(v), _d(p))
        return n.value

    def x_momentum(self, reg, g, dk, re, us, vs, un, vn, dus):
        r = self._reg(reg)
        self.L.xmomentum_(self._ir(reg.nx), self._ir(reg.ny), *r, self._dr(dk), self._dr(re),
                          _d(reg.dPRporos), _d(reg.dPRporc1), _d(reg.dPRporc2),
                          *[_d(g[nm]) for nm in ("rbn", "rgn", "rac", "rbc", "dju", "xec", "yec", "xzn", "yzn",
                                                 "xeu", "yeu", "xzu", "yzu")],
                          _d(us), _d(vs), _d(un), _d(vn), _d(dus))

    def y_momentum(self, reg, g, dk, re, fr, d, dn, us, vs, un, vn, dvs):
        r = self._reg(reg)
        self.L.ymomentum_(self._ir(reg.nx), self._ir(reg.ny), *r, self._dr(dk), self._dr(re), self._dr(fr),
                          _d(reg.dPRporos), _d(reg.dPRporc1), _d(reg.dPRporc2),
                          *[_d(g[nm]) for nm in ("ran", "rbn", "rbc", "rgc", "djv", "xen", "yen", "xzc", "yzc",
                                                 "xev", "yev", "xzv", "yzv")],
                          _d(d), _d(dn), _d(us), _d(vs), _d(un), _d(vn), _d(dvs))

    def naux_momentum(self, reg, g, mqiter, dk, re, fr, qtol, d, dn, un, vn, us, vs):
        r = self._reg(reg)
        f = self.L.nauxmomentum_
        f.restype = C.c_int
        return f(self._ir(reg.nx), self._ir(reg.ny), self._ir(mqiter), *r,
                 self._dr(dk), self._dr(re), self._dr(fr), self._dr(qtol),
                 _d(reg.dPRporos), _d(reg.dPRporc1), _d(reg.dPRporc2), _d(reg.dBCVal),
                 *[_d(g[nm]) for nm in ("ran", "rbn", "rgn", "rac", "rbc", "rgc", "dju", "djv",
                                        "xec", "yec", "xzn", "yzn", "xen", "yen", "xzc", "yzc",
                                        "xeu", "yeu", "xzu", "yzu", "xev", "yev", "xzv", "yzv")],
                 _d(d), _d(dn), _d(un), _d(vn), _d(us), _d(vs))

    def temp_bound_cond(self, reg, th, t):
        self.L.tempboundcond_(self._ir(reg.nx), self._ir(reg.ny), _i(reg.nreg), _i(reg.nRegBrd),
                              _i(th.nTRgType), _i(th.nTemBdTp), _d(th.dTRgVal), _d(reg.dBCVal), _d(t))

    def therm_energy(self, reg, th, g, dk, pe, un, vn, u, v, tn, t):
        self.L.thermenergy_(self._ir(reg.nx), self._ir(reg.ny), _i(reg.nreg), _i(reg.nRegBrd),
                            _i(th.nTRgType), _i(th.nTemBdTp), self._dr(dk), self._dr(pe),
                            _d(th.dTRgVal), _d(th.dHGSTval), _d(reg.dBCVal),
                            *[_d(g[nm]) for nm in ("rau", "rbu", "rbv", "rgv", "djc", "xeu", "yeu", "xzv", "yzv",
                                                   "xec", "yec", "xzc", "yzc")],
                            _d(un), _d(vn), _d(u), _d(v), _d(tn), _d(t))

    def eq_state(self, nx, ny, uref, densref, tmax, tref, rconst, p, t, den):
        self.L.eqstate_(self._ir(nx), self._ir(ny), *[self._dr(x) for x in (uref, densref, tmax, tref, rconst)],
                        _d(p), _d(t), _d(den))

    def filter(self, reg, th, ncomp, fp, q):
        self.L.filter_(self._ir(reg.nx), self._ir(reg.ny), self._ir(ncomp), _i(reg.nreg), _i(reg.nRegBrd),
                       _i(reg.nRegType), _i(reg.nMomBdTp), _i(th.nTRgType), self._dr(fp), _d(q))

    def dmax_norm(self, nx, ny, u):
        return self.L.dmaxnorm_(self._ir(nx), self._ir(ny), _d(u))

    def diff_max_norm(self, nx, ny, un, u):
        return self.L.diffmaxnorm_(self._ir(nx), self._ir(ny), _d(un), _d(u))


def run_reference_program(rundir: str, exe: str | None = None, timeout: float | None = None,
                          taskset_core: int | None = None) -> str:
    """Run the reference executable in ``rundir`` (reads input.dat there); returns stdout."""
    import subprocess
    exe = exe or os.path.join(REF_DIR, "std", "wolfd2")
    cmd = [exe]
    if taskset_core is not None:
        cmd = ["taskset", "-c", str(taskset_core)] + cmd
    res = subprocess.run(cmd, cwd=rundir, capture_output=True, text=True, timeout=timeout)
    if res.returncode != 0:
        raise RuntimeError(f"reference run failed ({res.returncode}): {res.stdout[-2000:]} {res.stderr[-2000:]}")
    return res.stdout
