"""The reference's hot-path call interface, backed by the HIP library.

Each function mirrors one subroutine of the reference with its argument
meaning (momentum.f:nAuxMomentum/XMomentum/YMomentum/AltTridLU,
pressure.f:Ppe/Divergence, utility.f:Project/DMaxNorm/DiffMaxNorm,
bound_cond.f:VelBoundCond/PresBoundCond/VelOutflowBCs/TempBoundCond,
thermal.f:ThermEnergy/EqState, utility.f:Filter): fields are numpy
float64 arrays of shape (mny+1, mnx+1) indexed [j, i] (the reference's
(0:mnx,0:mny) column-major arrays), updated in place; region tables and
metrics come from :mod:`wolfd2_amd.case`.  Every call runs on the GPU through
the C ABI of include/wolfd2_amd.h; a bad boundary or solver flag raises
:class:`Wd2Error` where the reference would STOP.

:class:`Solver` is the device-resident driver: it keeps u, v, p in HBM and
advances whole time steps (main.f:690-972).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import case
from ._lib import AdjStats, Params, StepOptions, Wd2Error, check, dp, fp, ip, lib

__all__ = ["vel_bound_cond", "pres_bound_cond", "vel_outflow_bcs", "project", "divergence", "ppe",
           "x_momentum", "y_momentum", "naux_momentum", "alt_trid_lu", "dmax_norm", "diff_max_norm",
           "temp_bound_cond", "therm_energy", "eq_state", "filter",
           "Solver", "Ensemble", "Adjoint", "Wd2Error", "make_params"]


def _ld(a: np.ndarray) -> int:
    return a.shape[1]


def vel_bound_cond(reg: case.Regions, u: np.ndarray, v: np.ndarray) -> None:
    check(lib().wd2_vel_bound_cond(reg.nx, reg.ny, _ld(u), reg.mgri, reg.mgrj, ip(reg.nreg), ip(reg.nRegBrd),
                                   ip(reg.nMomBdTp), fp(reg.dBCVal), dp(u), dp(v)))


def vel_outflow_bcs(reg: case.Regions, u: np.ndarray, v: np.ndarray) -> None:
    check(lib().wd2_vel_outflow_bcs(reg.nx, reg.ny, _ld(u), reg.mgri, reg.mgrj, ip(reg.nreg), ip(reg.nRegBrd),
                                    ip(reg.nMomBdTp), fp(reg.dBCVal), dp(u), dp(v)))


def pres_bound_cond(reg: case.Regions, p: np.ndarray) -> None:
    check(lib().wd2_pres_bound_cond(reg.nx, reg.ny, _ld(p), reg.mgri, reg.mgrj, ip(reg.nreg), ip(reg.nRegBrd),
                                    ip(reg.nRegType), ip(reg.nMomBdTp), fp(reg.dBCVal), dp(p)))


def project(reg: case.Regions, g: case.Grid, dk: float, p, u, v) -> None:
    check(lib().wd2_project(reg.nx, reg.ny, g.ld, reg.mgri, reg.mgrj, ip(reg.nreg), ip(reg.nRegBrd),
                            ip(reg.nRegType), ip(reg.nMomBdTp), C.c_double(dk), dp(g.dju), dp(g.djv), dp(g.yeu),
                            dp(g.xzv), dp(g.yzu), dp(g.xev), dp(p), dp(u), dp(v)))


def temp_bound_cond(reg: case.Regions, th: case.Thermal, t: np.ndarray) -> None:
    check(lib().wd2_temp_bound_cond(reg.nx, reg.ny, _ld(t), reg.mgri, reg.mgrj, ip(reg.nreg), ip(reg.nRegBrd),
                                    ip(th.nTRgType), ip(th.nTemBdTp), fp(th.dTRgVal), fp(reg.dBCVal), dp(t)))


def therm_energy(reg: case.Regions, th: case.Thermal, g: case.Grid, dk: float, pe: float, un, vn, u, v, tn,
                 t) -> None:
    check(lib().wd2_therm_energy(reg.nx, reg.ny, g.ld, reg.mgri, reg.mgrj, ip(reg.nreg), ip(reg.nRegBrd),
                                 ip(th.nTRgType), ip(th.nTemBdTp), C.c_double(dk), C.c_double(pe),
                                 fp(th.dTRgVal), fp(th.dHGSTval), fp(reg.dBCVal),
                                 *[dp(g.metrics[k]) for k in ("rau", "rbu", "rbv", "rgv", "djc", "xeu", "yeu",
                                                              "xzv", "yzv", "xec", "yec", "xzc", "yzc")],
                                 dp(un), dp(vn), dp(u), dp(v), dp(tn), dp(t)))


def eq_state(g: case.Grid, uref: float, densref: float, tmax: float, tref: float, rconst: float, p, t,
             den) -> None:
    check(lib().wd2_eq_state(g.nx, g.ny, g.ld, *[C.c_double(x) for x in (uref, densref, tmax, tref, rconst)],
                             dp(p), dp(t), dp(den)))


def filter(reg: case.Regions, th: case.Thermal, ncomp: int, fpar: float, q: np.ndarray) -> None:
    check(lib().wd2_filter(reg.nx, reg.ny, _ld(q), reg.mgri, reg.mgrj, ncomp, ip(reg.nreg), ip(reg.nRegBrd),
                           ip(reg.nRegType), ip(reg.nMomBdTp), ip(th.nTRgType), C.c_double(fpar), dp(q)))


def divergence(g: case.Grid, nloc: int, u, v, div) -> None:
    if nloc == 1:
        ms = (g.xeu, g.yeu, g.xzv, g.yzv)
    else:
        ms = (g.xev, g.yev, g.xzu, g.yzu)
    check(lib().wd2_divergence(g.nx, g.ny, g.ld, nloc, *[dp(m) for m in ms], dp(u), dp(v), dp(div)))


def ppe(reg: case.Regions, g: case.Grid, prm: case.Params, u, v, p, solver=None, msorit=None, sortol=None,
        sorrel=None) -> int:
    """pressure.f:Ppe; returns nSorConv."""
    n = C.c_int(0)
    check(lib().wd2_ppe(reg.nx, reg.ny, g.ld, reg.mgri, reg.mgrj, ip(reg.nreg), ip(reg.nRegBrd), ip(reg.nRegType),
                        int(prm.cartesian), prm.nPpeSolver if solver is None else solver,
                        prm.msorit if msorit is None else msorit, C.byref(n), C.c_double(prm.dk),
                        C.c_double(prm.sortol if sortol is None else sortol),
                        C.c_double(prm.sorrel if sorrel is None else sorrel),
                        dp(g.rau), dp(g.rbu), dp(g.rbv), dp(g.rgv), dp(g.xeu), dp(g.yeu), dp(g.xzv), dp(g.yzv),
                        dp(u), dp(v), dp(p)))
    return n.value


def x_momentum(reg, g, prm, us, vs, un, vn, dus) -> None:
    check(lib().wd2_x_momentum(reg.nx, reg.ny, g.ld, reg.mgri, reg.mgrj, ip(reg.nreg), ip(reg.nRegBrd),
                               ip(reg.nRegType), ip(reg.nMomBdTp), C.c_double(prm.dk), C.c_double(prm.re),
                               fp(reg.dPRporos), fp(reg.dPRporc1), fp(reg.dPRporc2),
                               *[dp(g.metrics[k]) for k in ("rbn", "rgn", "rac", "rbc", "dju", "xec", "yec", "xzn",
                                                            "yzn", "xeu", "yeu", "xzu", "yzu")],
                               dp(us), dp(vs), dp(un), dp(vn), dp(dus)))


def y_momentum(reg, g, prm, d, dn, us, vs, un, vn, dvs) -> None:
    check(lib().wd2_y_momentum(reg.nx, reg.ny, g.ld, reg.mgri, reg.mgrj, ip(reg.nreg), ip(reg.nRegBrd),
                               ip(reg.nRegType), ip(reg.nMomBdTp), C.c_double(prm.dk), C.c_double(prm.re),
                               C.c_double(prm.fr), fp(reg.dPRporos), fp(reg.dPRporc1), fp(reg.dPRporc2),
                               *[dp(g.metrics[k]) for k in ("ran", "rbn", "rbc", "rgc", "djv", "xen", "yen", "xzc",
                                                            "yzc", "xev", "yev", "xzv", "yzv")],
                               dp(d), dp(dn), dp(us), dp(vs), dp(un), dp(vn), dp(dvs)))


def naux_momentum(reg, g, prm, d, dn, un, vn, us, vs, mqiter=None, qtol=None) -> int:
    """momentum.f:nAuxMomentum; returns the q-l iteration count or -1."""
    n = C.c_int(0)
    check(lib().wd2_naux_momentum(
        reg.nx, reg.ny, g.ld, reg.mgri, reg.mgrj, prm.mqiter if mqiter is None else mqiter,
        ip(reg.nreg), ip(reg.nRegBrd), ip(reg.nRegType), ip(reg.nMomBdTp),
        C.c_double(prm.dk), C.c_double(prm.re), C.c_double(prm.fr),
        C.c_double(prm.qtol if qtol is None else qtol),
        fp(reg.dPRporos), fp(reg.dPRporc1), fp(reg.dPRporc2), fp(reg.dBCVal),
        *[dp(g.metrics[k]) for k in ("ran", "rbn", "rgn", "rac", "rbc", "rgc", "dju", "djv", "xec", "yec", "xzn",
                                     "yzn", "xen", "yen", "xzc", "yzc", "xeu", "yeu", "xzu", "yzu", "xev", "yev",
                                     "xzv", "yzv")],
        dp(d), dp(dn), dp(un), dp(vn), dp(us), dp(vs), C.byref(n)))
    return n.value


def alt_trid_lu(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """momentum.f:AltTridLU on an (n, 3) row array (sub, diag, super); returns x."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    x = np.array(b, dtype=np.float64, copy=True)
    check(lib().wd2_alt_trid_lu(x.size, dp(a), dp(x)))
    return x


def dmax_norm(nx: int, ny: int, u: np.ndarray) -> float:
    out = C.c_double(0)
    check(lib().wd2_dmax_norm(nx, ny, _ld(u), dp(u), C.byref(out)))
    return out.value


def diff_max_norm(nx: int, ny: int, un: np.ndarray, u: np.ndarray) -> float:
    out = C.c_double(0)
    check(lib().wd2_diff_max_norm(nx, ny, _ld(u), dp(un), dp(u), C.byref(out)))
    return out.value


def make_params(prm: case.Params) -> Params:
    return Params(int(prm.cartesian), prm.mqiter, prm.nPpeSolver, prm.msorit, prm.qtol, prm.sortol, prm.sorrel,
                  prm.dk, prm.re, prm.fr)


class Solver:
    """Device-resident driver of main.f:600-972 for one case (wd2_ctx_*)."""

    def __init__(self, g: case.Grid, reg: case.Regions, prm: case.Params):
        self.g, self.reg, self.prm = g, reg, prm
        self.ctx = C.c_void_p()
        mets = (C.POINTER(C.c_double) * 30)(*[dp(g.metrics[k]) for k in case.METRIC_NAMES])
        self._mets = mets
        check(lib().wd2_ctx_create(C.byref(self.ctx), g.nx, g.ny, g.ld, reg.mgri, reg.mgrj, ip(reg.nreg),
                                   ip(reg.nRegBrd), ip(reg.nRegType), ip(reg.nMomBdTp), fp(reg.dBCVal),
                                   fp(reg.dPRporos), fp(reg.dPRporc1), fp(reg.dPRporc2), mets))
        self.params = make_params(prm)
        self.thermal = False
        for (ir, jr, side, var), values in reg.profiles.items():
            self.set_bc_profile(ir, jr, side, var, values)

    def set_bc_profile(self, ir: int, jr: int, side: int, var: int, values) -> None:
        """One value per grid line in place of the constant boundary value of
        variable ``var`` (case.VAR_U/V/P) on side ``side`` of region (ir, jr):
        indexed by j = 0..ny+1 on the west/east sides, by i = 0..nx+1 on the
        south/north sides.  None restores the constant."""
        if values is None:
            check(lib().wd2_ctx_set_bc_profile(self.ctx, ir, jr, side, var, None, 0))
            return
        a = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().wd2_ctx_set_bc_profile(self.ctx, ir, jr, side, var, dp(a), a.size))

    def close(self):
        if self.ctx:
            lib().wd2_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, prm: case.Params):
        self.prm = prm
        self.params = make_params(prm)

    def set_options(self, opt: case.StepOptions):
        """Energy equation, buoyancy, momentum/energy iterations and filters
        of the following steps (main.f:735-895)."""
        o = StepOptions(*[getattr(opt, f) for f, _ in StepOptions._fields_])
        check(lib().wd2_ctx_set_options(self.ctx, C.byref(o)))
        self.thermal = opt.nthermen == 1

    def set_thermal(self, th: case.Thermal):
        check(lib().wd2_ctx_set_thermal(self.ctx, ip(th.nTRgType), ip(th.nTemBdTp), fp(th.dTRgVal),
                                        fp(th.dHGSTval)))

    def set_thermal_fields(self, t=None, d=None):
        check(lib().wd2_ctx_set_thermal_fields(self.ctx, self.g.ld, dp(t), dp(d)))

    def get_thermal_fields(self):
        g = self.g
        t, d = np.zeros((g.ncol, g.ld)), np.zeros((g.ncol, g.ld))
        check(lib().wd2_ctx_get_thermal_fields(self.ctx, g.ld, dp(t), dp(d)))
        return t, d

    def set_precision(self, precision: str):
        """"fp64" (the reference's arithmetic, bit for bit) or "mixed" (FP64
        fields; FP32 momentum stencils and ADI chains; red-black SOR as FP64
        defect correction around FP32 sweeps)."""
        check(lib().wd2_ctx_set_precision(self.ctx, {"fp64": 0, "mixed": 1}[precision]))

    def set_graph(self, enable: bool = True) -> None:
        """Step as hipGraph replays where the setup allows (wd2_ctx_set_graph)."""
        check(lib().wd2_ctx_set_graph(self.ctx, 1 if enable else 0))

    def graph_stats(self) -> dict:
        a, b = C.c_long(), C.c_long()
        check(lib().wd2_ctx_graph_stats(self.ctx, C.byref(a), C.byref(b)))
        return {"replayed": a.value, "redone": b.value}

    def mg_autotune(self, cycles: int = 5) -> dict:
        """Pick the multigrid solver's fastest kernel shapes for this grid and
        precision (wd2_ctx_mg_autotune); returns the shape and its V-cycle time."""
        us = C.c_double(0)
        check(lib().wd2_ctx_mg_autotune(self.ctx, C.c_double(self.prm.dk), cycles, C.byref(us)))
        return {**self.mg_tune(), "vcycle_us": us.value}

    def mg_tune(self) -> dict:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().wd2_ctx_mg_get_tune(self.ctx, C.byref(a), C.byref(b), C.byref(c)))
        return {"nt0": a.value, "stage": b.value, "bside": c.value}

    def set_mg_tune(self, nt0: int, stage: int, bside: int) -> None:
        check(lib().wd2_ctx_mg_set_tune(self.ctx, nt0, stage, bside))

    def trid_autotune(self, reps: int = 20) -> dict:
        """Pick the mixed path's fastest window shape (lanes per window: 128,
        256 or 512) on the chains of the last q-l iteration
        (wd2_ctx_trid_autotune); returns it and the three shapes' times."""
        wt = C.c_int()
        us = (C.c_double * 3)()
        check(lib().wd2_ctx_trid_autotune(self.ctx, reps, C.byref(wt), us))
        return {"window_threads": wt.value, "us": {128: us[0], 256: us[1], 512: us[2]}}

    def set_trid_window(self, window_threads: int) -> None:
        check(lib().wd2_ctx_trid_set_window(self.ctx, window_threads))

    def set_mg_extrapolation(self, order: int = 1) -> int:
        """Start each multigrid solve from the pressure extrapolated from the
        steps before (wd2_ctx_set_mg_extrapolation): order 1, 2 p^n - p^(n-1);
        2, 3 p^n - 3 p^(n-1) + p^(n-2); 0 (or False), the plain start.
        Returns the previous order."""
        old = lib().wd2_ctx_set_mg_extrapolation(self.ctx, int(order))
        if old < 0:
            check(old)
        return old

    def mg_cycle_time(self, cycles: int = 5) -> float:
        us = C.c_double(0)
        check(lib().wd2_ctx_mg_cycle_time(self.ctx, C.c_double(self.prm.dk), cycles, C.byref(us)))
        return us.value

    def set_fields(self, u, v, p):
        check(lib().wd2_ctx_set_fields(self.ctx, self.g.ld, dp(u), dp(v), dp(p)))

    def get_fields(self):
        g = self.g
        u, v, p = (np.zeros((g.ncol, g.ld)) for _ in range(3))
        check(lib().wd2_ctx_get_fields(self.ctx, g.ld, dp(u), dp(v), dp(p)))
        return u, v, p

    def cold_start(self) -> int:
        n = C.c_int(0)
        check(lib().wd2_ctx_cold_start(self.ctx, C.byref(self.params), C.byref(n)))
        return n.value

    def step(self, nsteps: int = 1) -> np.ndarray:
        """Advance nsteps; returns rows (ql_iter, ppe_iter, |p1-pn|, |u1-un|,
        |v1-vn|), and |t1-tn| as a sixth column with the energy equation on."""
        nd = 4 if self.thermal else 3
        ql = (C.c_int * nsteps)()
        pi = (C.c_int * nsteps)()
        dif = (C.c_double * (nd * nsteps))()
        fn = lib().wd2_ctx_step4 if self.thermal else lib().wd2_ctx_step
        check(fn(self.ctx, C.byref(self.params), nsteps, ql, pi, dif))
        out = np.zeros((nsteps, 2 + nd))
        out[:, 0] = np.frombuffer(ql, dtype=np.int32)
        out[:, 1] = np.frombuffer(pi, dtype=np.int32)
        out[:, 2:] = np.frombuffer(dif, dtype=np.float64).reshape(nsteps, nd)
        return out


class Ensemble:
    """Independent cases advanced together (wd2_ens_*): each member's step is
    queued on its own stream before any is waited for.  Members are Solver
    objects, which keep their own fields, parameters and precision."""

    def __init__(self, members):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.ctx.value for m in self.members])
        self.ens = C.c_void_p()
        check(lib().wd2_ens_create(C.byref(self.ens), len(self.members), arr))

    def close(self):
        if self.ens:
            lib().wd2_ens_destroy(self.ens)
            self.ens = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps: int = 1) -> list:
        """Advance every member nsteps; returns one (nsteps, 5) table per
        member: ql_iter, ppe_iter, |p1-pn|, |u1-un|, |v1-vn|."""
        n = len(self.members)
        prm = (Params * n)(*[m.params for m in self.members])
        ql = (C.c_int * (n * nsteps))()
        pi = (C.c_int * (n * nsteps))()
        dif = (C.c_double * (3 * n * nsteps))()
        check(lib().wd2_ens_step(self.ens, prm, nsteps, ql, pi, dif))
        q = np.frombuffer(ql, dtype=np.int32).reshape(nsteps, n)
        pp = np.frombuffer(pi, dtype=np.int32).reshape(nsteps, n)
        d = np.frombuffer(dif, dtype=np.float64).reshape(nsteps, n, 3)
        out = []
        for i in range(n):
            t = np.zeros((nsteps, 5))
            t[:, 0], t[:, 1], t[:, 2:] = q[:, i], pp[:, i], d[:, i, :]
            out.append(t)
        return out



class Adjoint:
    """Reverse-mode gradient of a run of a Solver's time steps (wd2_adj_*):
    ``forward(nsteps)`` advances the solver from its current fields saving a
    checkpoint every ``checkpoint_every`` steps; ``backward(u_bar, v_bar,
    p_bar)`` returns the adjoint of the run's initial u, v, p and of the
    boundary values, given the adjoint of its final fields.  The case must be
    a one-region cavity with no-slip walls, multigrid Poisson solves, FP64;
    the gradient is that of the converged q-l loops and Poisson solves."""

    def __init__(self, solver: Solver, checkpoint_every: int = 20):
        self.solver = solver
        self.adj = C.c_void_p()
        check(lib().wd2_adj_create(C.byref(self.adj), solver.ctx, checkpoint_every))

    def close(self):
        if self.adj:
            lib().wd2_adj_destroy(self.adj)
            self.adj = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tolerances(self, fixed_point_tol: float = 0.0, max_iter: int = 0, poisson_tol: float = 0.0) -> None:
        check(lib().wd2_adj_set_tolerances(self.adj, C.c_double(fixed_point_tol), max_iter,
                                           C.c_double(poisson_tol)))

    def forward(self, nsteps: int) -> np.ndarray:
        """Advance nsteps with checkpoints; rows (ql_iter, ppe_iter)."""
        ql = (C.c_int * nsteps)()
        pi = (C.c_int * nsteps)()
        check(lib().wd2_adj_forward(self.adj, C.byref(self.solver.params), nsteps, ql, pi))
        return np.stack([np.frombuffer(ql, dtype=np.int32), np.frombuffer(pi, dtype=np.int32)], axis=1)

    def backward(self, u_bar, v_bar, p_bar=None) -> dict:
        """Adjoints of the initial fields (u0_bar, v0_bar, p0_bar: (ncol, ld)
        arrays), of the boundary values (bc_bar[side-1, var-1, k] for side
        W, E, S, N, var u, v, grid line k) and the pass's statistics."""
        g = self.solver.g
        out = [np.zeros((g.ncol, g.ld)) for _ in range(3)]
        W = max(g.nx, g.ny) + 2
        bc = np.zeros(8 * W)
        st = AdjStats()
        ub = np.ascontiguousarray(u_bar, dtype=np.float64)
        vb = np.ascontiguousarray(v_bar, dtype=np.float64)
        pb = None if p_bar is None else np.ascontiguousarray(p_bar, dtype=np.float64)
        check(lib().wd2_adj_backward(self.adj, g.ld, dp(ub), dp(vb), dp(pb), dp(out[0]), dp(out[1]), dp(out[2]),
                                     dp(bc), C.byref(st)))
        stats = {f: getattr(st, f) for f, _ in AdjStats._fields_}
        return {"u0_bar": out[0], "v0_bar": out[1], "p0_bar": out[2], "bc_bar": bc.reshape(4, 2, W),
                "stats": stats}

    def tangent(self, du0=None, dv0=None, dp0=None, dbc=None) -> dict:
        """Tangent-linear derivative of the last forward run along a direction
        of its initial fields (du0, dv0, dp0: (ncol, ld) arrays or None) and
        boundary values (dbc[side-1, var-1, k] as bc_bar, or None): the
        derivatives of the final u, v, p (duN, dvN, dpN) and the pass's
        statistics.  The run is repeated from its initial state."""
        g = self.solver.g
        out = [np.zeros((g.ncol, g.ld)) for _ in range(3)]
        W = max(g.nx, g.ny) + 2
        st = AdjStats()
        arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        d0 = [arr(du0), arr(dv0), arr(dp0)]
        bc = None if dbc is None else np.ascontiguousarray(np.asarray(dbc, dtype=np.float64).reshape(8 * W))
        check(lib().wd2_adj_tangent(self.adj, g.ld, dp(d0[0]), dp(d0[1]), dp(d0[2]), dp(bc), dp(out[0]),
                                    dp(out[1]), dp(out[2]), C.byref(st)))
        stats = {f: getattr(st, f) for f, _ in AdjStats._fields_}
        return {"duN": out[0], "dvN": out[1], "dpN": out[2], "stats": stats}
