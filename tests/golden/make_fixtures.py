"""Generate the golden fixtures under tests/golden/ from the reference solver.

Run in the build container (where /root/reference exists and `make -C oracle
ref` has built oracle/_ref):

    python tests/golden/make_fixtures.py [--only NAME ...]

Whole-run fixtures run the reference executable (oracle/_ref/std/wolfd2) on
decks written by decks.py and keep the final restart fields (p, u, v with
ghosts) and the per-step PrintDiff table.  Kernel fixtures call the
reference's own compiled subroutines (oracle/_ref/std/libwolfd2ref.so) on
seeded inputs over a stretched, skewed 17x11-node grid with a 2x2 region
layout, and keep inputs and outputs.  Everything is written as compressed
.npz (no pickles).
"""
from __future__ import annotations

import argparse
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import refdecks as decks  # noqa: E402
from pyoracle import Reference, run_reference_program  # noqa: E402
from wolfd2_amd import case  # noqa: E402

RUNS = {
    # name: (nodes nx, ny, Lx, Ly, deck kwargs)
    "c1_rbsor_64": (65, 65, 1.0, 1.0, dict(nts=1000, dt=1e-3, mu=0.01, ppe_solver="rb_sor")),
    "c1_sor_64": (65, 65, 1.0, 1.0, dict(nts=1000, dt=1e-3, mu=0.01, ppe_solver="sor")),
    "c1_tierb_64": (65, 65, 1.0, 1.0, dict(nts=200, dt=1e-3, mu=0.01, ppe_solver="rb_sor",
                                             ql_tolerance=1e-13, max_ql_iter=200, sor_relaxation=1.9,
                                             sor_tolerance=1e-13, max_sor_iter=200000)),
    "c1_lsor_32": (33, 33, 1.0, 1.0, dict(nts=100, dt=1e-3, mu=0.01, ppe_solver="lsor")),
    "c1_rblsor_32": (33, 33, 1.0, 1.0, dict(nts=100, dt=1e-3, mu=0.01, ppe_solver="rb_lsor")),
    "c1_parrblsor_32": (33, 33, 1.0, 1.0, dict(nts=100, dt=1e-3, mu=0.01, ppe_solver="par_rb_lsor",
                                                 sor_relaxation=1.2)),
    "c1_parrbsor_32": (33, 33, 1.0, 1.0, dict(nts=100, dt=1e-3, mu=0.01, ppe_solver="par_rb_sor",
                                                sor_relaxation=1.5)),
    # mass-conserving outlet; the fully-developed outlet diverges in the
    # reference after ~25 steps, so it is pinned on its first 15 steps only.
    "chan_64x16": (65, 17, 4.0, 1.0, dict(nts=200, dt=1e-3, mu=0.01, ppe_solver="rb_sor",
                                          bc_lines=["inlet 1 1 west normal_vel 1.0",
                                                    "outlet 1 1 east mass_cons"])),
    "chan_fd_64x16": (65, 17, 4.0, 1.0, dict(nts=15, dt=1e-3, mu=0.01, ppe_solver="rb_sor",
                                             bc_lines=["inlet 1 1 west normal_vel 1.0",
                                                       "outlet 1 1 east fully_dev press 0.1"])),
    "c5_re1000_128": (129, 129, 1.0, 1.0, dict(nts=50, dt=2e-4, mu=1e-3, ppe_solver="rb_sor")),
    "skew_rbsor_24x16": (25, 17, None, None, dict(nts=40, dt=1e-3, mu=0.01, ppe_solver="rb_sor",
                                                  cartesian=False)),
}


# Thermal runs (thermal_energy, buoyancy, filters, momentum/energy
# iterations): the drop-in link's parity cases.  Kept out of run_* because
# the device-resident context and the oracle's step are the cold-flow path.
THERMAL_RUNS = {
    "therm_cavity_32": (33, 33, 1.0, 1.0, dict(
        nts=60, dt=1e-3, mu=0.01, ppe_solver="rb_sor", nreg=(2, 1), i_borders=(17,), gas_constant=287.0,
        extra_input=["thermal_energy", "buoyancy", "max_me_iter 3", "me_tolerance 1e-9", "filter_u 200.0",
                     "filter_v 200.0", "filter_t 50.0", "ref_temperature 300.0", "max_ref_temperature 310.0"],
        bc_lines=["wall 1 1 west temper 310.0", "wall 2 1 east temper 300.0", "wall 1 1 north tangent_vel 0.5",
                  "heat_generation_region 2 1 40.0"])),
}


def thermal_run_fixture(name: str, out_dir: str) -> None:
    nx, ny, lx, ly, kw = THERMAL_RUNS[name]
    x, y = case.uniform_nodes(nx, ny, lx, ly)
    with tempfile.TemporaryDirectory(prefix=f"wd2fx_{name}_") as rd:
        d = decks.Deck(**kw)
        d.write(rd, x, y)
        stdout = run_reference_program(rd)
        rst = decks.read_restart(os.path.join(rd, "out.rst"))
    lines = [ln for ln in stdout.splitlines() if decks._DIFF_RE.match(ln)]
    sl = np.s_[: ny + 2, : nx + 2]
    np.savez_compressed(
        os.path.join(out_dir, f"thermrun_{name}.npz"),
        x=x, y=y, nts=kw["nts"], dt=kw["dt"], mu=kw["mu"], rho=1.0, gas_constant=kw["gas_constant"],
        deck=d.input_text(),
        k=rst["k"], dtime=rst["dtime"], diff_lines="\n".join(lines),
        p=rst["p"][sl], u=rst["u"][sl], v=rst["v"][sl], t=rst["t"][sl],
    )
    print(f"{name}: steps={rst['k']} rows={len(lines)} last: {lines[-1].strip()}")


def run_fixture(name: str, out_dir: str) -> None:
    nx, ny, lx, ly, kw = RUNS[name]
    if lx is None:
        x, y = case.skewed_nodes(nx, ny, seed=7)
    else:
        x, y = case.uniform_nodes(nx, ny, lx, ly)
    with tempfile.TemporaryDirectory(prefix=f"wd2fx_{name}_") as rd:
        d = decks.Deck(**kw)
        d.write(rd, x, y)
        stdout = run_reference_program(rd)
        rst = decks.read_restart(os.path.join(rd, "out.rst"))
        table = decks.parse_print_diff(stdout)
    sl = np.s_[: ny + 2, : nx + 2]
    np.savez_compressed(
        os.path.join(out_dir, f"run_{name}.npz"),
        x=x, y=y, nts=kw["nts"], dt=kw["dt"], mu=kw["mu"], rho=1.0,
        ppe_solver=case.PPE_SOLVERS[kw["ppe_solver"]],
        deck=d.input_text(), k=rst["k"], dtime=rst["dtime"],
        p=rst["p"][sl], u=rst["u"][sl], v=rst["v"][sl], table=table,
    )
    print(f"{name}: steps={rst['k']} table={table.shape} ql[-1]={table[-1, 2]:.0f} ppe[-1]={table[-1, 3]:.0f}")


def kernel_case(seed: int = 11):
    """The 17x11 skewed grid, 2x2 regions with every boundary type, a porous region."""
    nx, ny = 17, 11
    x, y = case.skewed_nodes(nx, ny, seed=3)
    reg = case.make_regions(
        nx, ny, nreg=(2, 2), i_borders=(8,), j_borders=(5,),
        walls=[(1, 2, "north", "no_slip", 1.0, 0.1), (2, 1, "south", "no_stress", 0.0, 0.2)],
        inlets=[(1, 1, "west", 0.7, 0.1, 0.05)],
        outlets=[(2, 2, "east", "mass_cons", 0.0, 0.0, 0.3), (2, 1, "east", "fully_dev", 0.01, 0.02, 0.4)],
        porous=[(2, 2, 0.6, 0.8, 0.3)])
    return nx, ny, x, y, reg


def blockage_case():
    """3x3 regions with a blockage in the middle (the PPE/momentum blockage rows)."""
    nx, ny = 19, 13
    x, y = case.uniform_nodes(nx, ny)
    reg = case.make_regions(nx, ny, nreg=(3, 3), i_borders=(7, 13), j_borders=(5, 9),
                            walls=[(2, 3, "north", "no_slip", 1.0, 0.0)], blockages=[(2, 2)])
    return nx, ny, x, y, reg


def kernel_fixtures(out_dir: str) -> None:
    for tag, builder in (("skew", kernel_case), ("block", blockage_case)):
        nx, ny, x, y, reg = builder()
        R = Reference.for_grid((tag, nx, ny))
        gx, gy = R.field(), R.field()
        gx[1:ny + 1, 1:nx + 1] = x
        gy[1:ny + 1, 1:nx + 1] = y
        g = R.grid(nx, ny, gx, gy, {})
        rng = np.random.default_rng(2024)
        sl = np.s_[: ny + 2, : nx + 2]

        def rf(scale=1.0):
            f = R.field()
            f[sl] = scale * rng.standard_normal((ny + 2, nx + 2))
            return f

        out = dict(nx=nx, ny=ny, x=x, y=y)
        for nm in ("nreg", "nRegBrd", "nRegType", "nMomBdTp", "dBCVal", "dPRporos", "dPRporc1", "dPRporc2"):
            out[f"reg_{nm}"] = getattr(reg, nm)
        u, v = rf(), rf()
        out["vbc_u_in"], out["vbc_v_in"] = u[sl].copy(), v[sl].copy()
        R.vel_bound_cond(reg, u, v)
        out["vbc_u_out"], out["vbc_v_out"] = u[sl].copy(), v[sl].copy()
        u, v = rf(), rf()
        out["vob_u_in"], out["vob_v_in"] = u[sl].copy(), v[sl].copy()
        R.vel_outflow_bcs(reg, u, v)
        out["vob_u_out"], out["vob_v_out"] = u[sl].copy(), v[sl].copy()
        p = rf()
        out["pbc_p_in"] = p[sl].copy()
        R.pres_bound_cond(reg, p)
        out["pbc_p_out"] = p[sl].copy()
        u, v, p = rf(), rf(), rf()
        out["prj_u_in"], out["prj_v_in"], out["prj_p"] = u[sl].copy(), v[sl].copy(), p[sl].copy()
        out["prj_dk"] = 0.01
        R.project(reg, g, 0.01, p, u, v)
        out["prj_u_out"], out["prj_v_out"] = u[sl].copy(), v[sl].copy()
        div = R.field()
        R.divergence(nx, ny, 1, g, u, v, div)
        out["div_u"], out["div_v"], out["div_out"] = u[sl].copy(), v[sl].copy(), div[sl].copy()
        for solver in range(1, 7):
            for cart in (1, 0):
                u, v, p = rf(0.1), rf(0.1), rf()
                key = f"ppe{solver}c{cart}"
                out[f"{key}_u"], out[f"{key}_v"], out[f"{key}_p_in"] = u[sl].copy(), v[sl].copy(), p[sl].copy()
                n = R.ppe(reg, g, bool(cart), solver, 25, 0.01, 1e-8, 1.3, u, v, p)
                out[f"{key}_p_out"], out[f"{key}_n"] = p[sl].copy(), n
        us, vs, un, vn = rf(0.3), rf(0.3), rf(0.3), rf(0.3)
        for nm, a in (("us", us), ("vs", vs), ("un", un), ("vn", vn)):
            out[f"mom_{nm}"] = a[sl].copy()
        dus, dvs, dd, dn = R.field(), R.field(), R.field(), R.field()
        R.x_momentum(reg, g, 0.01, 100.0, us, vs, un, vn, dus)
        R.y_momentum(reg, g, 0.01, 100.0, 1.0 / 9.81, dd, dn, us, vs, un, vn, dvs)
        out["mom_dus"], out["mom_dvs"] = dus[sl].copy(), dvs[sl].copy()
        for mq, qt in ((1, 1e-4), (30, 1e-10)):
            us2, vs2 = R.field(), R.field()
            n = R.naux_momentum(reg, g, mq, 0.01, 100.0, 1.0 / 9.81, qt, dd, dn, un, vn, us2, vs2)
            out[f"naux{mq}_n"], out[f"naux{mq}_us"], out[f"naux{mq}_vs"] = n, us2[sl].copy(), vs2[sl].copy()
        out["dmax"] = R.dmax_norm(nx, ny, dus)
        out["diffmax"] = R.diff_max_norm(nx, ny, us, vs)
        a = rng.standard_normal((301, 3))
        a[:, 1] += 4.0
        b = rng.standard_normal(301)
        out["trid_a"], out["trid_b"] = a.copy(), b.copy()
        out["trid_x"] = R.alt_trid_lu(a, b)[1]
        np.savez_compressed(os.path.join(out_dir, f"kernels_{tag}.npz"), **out)
        print(f"kernels_{tag}: {len(out)} arrays; naux n = {out['naux1_n']}, {out['naux30_n']}")


def thermal_layouts():
    """The two kernel grids with thermal tables: heat generation, a fixed-
    temperature region, fixed-temperature and heat-flux walls."""
    nx, ny, x, y, reg = kernel_case()
    th = case.make_thermal(reg, heat=[(2, 1, 0.5)], fixed=[(1, 2, 0.3)],
                           walls=[(1, 1, "west", "temper", 1.0), (2, 2, "north", "flux", 0.05)])
    yield "skew", nx, ny, x, y, reg, th
    nx, ny, x, y, reg = blockage_case()
    th = case.make_thermal(reg, heat=[(3, 3, 0.8)], fixed=[(1, 3, -0.2)],
                           walls=[(2, 2, sd, "flux", 0.0) for sd in ("west", "east", "south", "north")]
                           + [(1, 1, "south", "temper", 1.0), (3, 1, "east", "temper", 0.0)])
    yield "block", nx, ny, x, y, reg, th


def thermal_fixtures(out_dir: str) -> None:
    """TempBoundCond, ThermEnergy, EqState and Filter (u, v, t) of the
    reference on seeded fields, as thermal_<layout>.npz."""
    for tag, nx, ny, x, y, reg, th in thermal_layouts():
        R = Reference.for_grid(("thermal", tag, nx, ny))
        gx, gy = R.field(), R.field()
        gx[1:ny + 1, 1:nx + 1] = x
        gy[1:ny + 1, 1:nx + 1] = y
        g = R.grid(nx, ny, gx, gy, {})
        rng = np.random.default_rng(4242)
        sl = np.s_[: ny + 2, : nx + 2]

        def rf(scale=1.0):
            f = R.field()
            f[sl] = scale * rng.standard_normal((ny + 2, nx + 2))
            return f

        out = dict(nx=nx, ny=ny, x=x, y=y)
        for nm in ("nreg", "nRegBrd", "nRegType", "nMomBdTp", "dBCVal", "dPRporos", "dPRporc1", "dPRporc2"):
            out[f"reg_{nm}"] = getattr(reg, nm)
        for nm in ("nTRgType", "nTemBdTp", "dTRgVal", "dHGSTval"):
            out[f"th_{nm}"] = getattr(th, nm)
        t = rf()
        out["tbc_in"] = t[sl].copy()
        R.temp_bound_cond(reg, th, t)
        out["tbc_out"] = t[sl].copy()
        un, vn, u, v, tn, t = rf(0.3), rf(0.3), rf(0.3), rf(0.3), rf(), rf()
        for nm, a in (("un", un), ("vn", vn), ("u", u), ("v", v), ("tn", tn), ("t", t)):
            out[f"te_{nm}"] = a[sl].copy()
        out["te_dk"], out["te_pe"] = 0.01, 70.0
        R.therm_energy(reg, th, g, 0.01, 70.0, un, vn, u, v, tn, t)
        out["te_out"] = t[sl].copy()
        p, t, den = rf(0.1), rf(0.5), R.field()
        out["eqs_p"], out["eqs_t"] = p[sl].copy(), t[sl].copy()
        out["eqs_args"] = np.array([1.5, 1.2, 400.0, 300.0, 287.0])
        R.eq_state(nx, ny, *out["eqs_args"], p, t, den)
        out["eqs_out"] = den[sl].copy()
        for comp, nm in ((1, "u"), (2, "v"), (4, "t")):
            q = rf()
            out[f"filt_{nm}_in"] = q[sl].copy()
            R.filter(reg, th, comp, 20.0, q)
            out[f"filt_{nm}_out"] = q[sl].copy()
        np.savez_compressed(os.path.join(out_dir, f"thermal_{tag}.npz"), **out)
        print(f"thermal_{tag}: {len(out)} arrays")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    ap.add_argument("--out", default=HERE)
    args = ap.parse_args()
    names = args.only or (["kernels", "thermal"] + list(RUNS) + list(THERMAL_RUNS))
    for nm in names:
        if nm == "kernels":
            kernel_fixtures(args.out)
        elif nm == "thermal":
            thermal_fixtures(args.out)
        elif nm in THERMAL_RUNS:
            thermal_run_fixture(nm, args.out)
        else:
            run_fixture(nm, args.out)


if __name__ == "__main__":
    main()
