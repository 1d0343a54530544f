"""The tangent-linear derivative of a run of time steps (wd2_adj_tangent, no
counterpart in the reference) against the adjoint and central differences.

The tangent-linear pass runs the forward map again with each step followed by
its derivative along a direction d of the initial u, v, p and the boundary
values; the adjoint gives the gradient of a functional J(final fields) =
<w, fields>.  Both are exact derivatives of the same map (converged q-l loops
and Poisson solves) computed by different code: the tangent-linear pass
evaluates the forward assembly on dual numbers and solves the linearised q-l
iteration forwards, the adjoint transposes every piece by hand.  So
<w, J d> = <J^T w, d> (the dot-product test) holds to the solvers'
tolerances, and a wrong transpose, sign or index in either shows as an O(1)
difference; central differences of the forward run pin both.
"""
import numpy as np
import pytest

from wolfd2_amd import api, autograd, case, workloads

pytestmark = pytest.mark.gpu


def _stretched(nx, ny):
    s = np.arange(nx) / (nx - 1.0)
    t = np.arange(ny) / (ny - 1.0)
    xs = s + 0.12 * np.sin(np.pi * s) * s
    yt = t - 0.08 * np.sin(2 * np.pi * t)
    return np.broadcast_to(xs[None, :], (ny, nx)).copy(), np.broadcast_to(yt[:, None], (ny, nx)).copy()


def _small_case():
    x, y = _stretched(34, 30)
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(g.nx, g.ny)
    prm = case.Params.from_physical(2e-3, 1.0, 1.0 / 100.0, nPpeSolver=7, msorit=100, sortol=1e-15, qtol=1e-15,
                                    mqiter=80)
    return g, reg, prm, x, y


@pytest.mark.parametrize("nsteps,every", [(1, 1), (5, 2)])
def test_tangent_linear_dot_product_and_central_differences_small(nsteps, every):
    """Stretched 33x29 cavity from a developed state: on random directions of
    the lid, the initial u, v (updated points) and p, <w, J d> from the
    tangent-linear pass equals <J^T w, d> from the adjoint to 1e-9 relative
    (measured 4e-14 ... 9e-11, both solves at their default tolerances) and
    central differences of the forward run to 1e-7 (after one step the lid
    direction gives exactly zero in all three)."""
    g, reg, prm, x, y = _small_case()
    lid = workloads.lid_modes(x, y, reg, seed=5)
    sol = api.Solver(g, reg, prm)
    sol.set_bc_profile(1, 1, case.NORTH, case.VAR_U, lid)
    z = np.zeros((g.ncol, g.ld))
    sol.set_fields(z, z, z)
    sol.step(4)
    u0, v0, p0 = sol.get_fields()
    rng = np.random.default_rng(100 + nsteps)
    su, sv = workloads.interior_u(g), workloads.interior_v(g)
    sp = np.s_[2:g.ny + 1, 2:g.nx + 1]
    mask_u, mask_v, mask_p = (np.zeros((g.ncol, g.ld)) for _ in range(3))
    mask_u[su] = 1
    mask_v[sv] = 1
    mask_p[sp] = 1
    wu, wv, wp = (rng.standard_normal((g.ncol, g.ld)) for _ in range(3))
    wu *= mask_u
    wv *= mask_v
    wp *= mask_p
    adj = api.Adjoint(sol, every)
    W = max(g.nx, g.ny) + 2

    def J(lid_, u_, v_, p_, with_tape=False):
        sol.set_bc_profile(1, 1, case.NORTH, case.VAR_U, lid_)
        sol.set_fields(u_, v_, p_)
        tab = adj.forward(nsteps) if with_tape else sol.step(nsteps)[:, :2]
        assert np.all(tab[:, 0] > 0)
        u, v, p = sol.get_fields()
        return float(np.sum(wu * u) + np.sum(wv * v) + np.sum(wp * p))

    J(lid, u0, v0, p0, with_tape=True)
    r = adj.backward(wu, wv, wp)
    assert r["stats"]["adj_unconverged"] == 0
    gl = r["bc_bar"][case.NORTH - 1, case.VAR_U - 1, 2:g.nx + 1]
    worst_dot = worst_fd = 0.0
    for k in range(4):
        dl = np.zeros(g.nx + 2)
        dl[2:g.nx + 1] = rng.standard_normal(g.nx - 1)
        du, dv, dpp = (rng.standard_normal((g.ncol, g.ld)) for _ in range(3))
        du *= mask_u
        dv *= mask_v
        dpp *= mask_p
        which = k % 4   # lid, velocity, pressure, everything
        if which == 0:
            du[:] = dv[:] = dpp[:] = 0.0
        elif which == 1:
            dl[:] = 0.0
            dpp[:] = 0.0
        elif which == 2:
            dl[:] = 0.0
            du[:] = dv[:] = 0.0
        dbc = np.zeros((4, 2, W))
        dbc[case.NORTH - 1, case.VAR_U - 1, :g.nx + 2] = dl
        # the tangent-linear pass re-runs the taped forward run with the
        # boundary values in force: the run's lid (the differences below
        # leave a perturbed one)
        sol.set_bc_profile(1, 1, case.NORTH, case.VAR_U, lid)
        t = adj.tangent(du, dv, dpp, dbc)
        assert t["stats"]["adj_unconverged"] == 0
        tl = float(np.sum(wu * t["duN"]) + np.sum(wv * t["dvN"]) + np.sum(wp * t["dpN"]))
        ad = float(np.sum(gl * dl[2:g.nx + 1]) + np.sum(r["u0_bar"] * du) + np.sum(r["v0_bar"] * dv) +
                   np.sum(r["p0_bar"] * dpp))
        eps = 1e-6
        fd = (J(lid + eps * dl, u0 + eps * du, v0 + eps * dv, p0 + eps * dpp) -
              J(lid - eps * dl, u0 - eps * du, v0 - eps * dv, p0 - eps * dpp)) / (2 * eps)
        if ad == 0.0 and fd == 0.0:
            # one step: a new lid value reaches only the final ghosts, so the
            # interior fields do not depend on it; all three are exactly zero
            rel_dot = rel_fd = abs(tl)
        else:
            rel_dot = abs(tl - ad) / abs(ad)
            rel_fd = abs(tl - fd) / abs(fd)
        worst_dot = max(worst_dot, rel_dot)
        worst_fd = max(worst_fd, rel_fd)
        print(f"    direction {k} ({['lid', 'velocity', 'pressure', 'all'][which]}): tangent {tl:.14e} "
              f"adjoint {ad:.14e} FD {fd:.10e}; |TL-AD|/|AD| {rel_dot:.2e}, |TL-FD|/|FD| {rel_fd:.2e}; "
              f"TL q-l iterations max {t['stats']['adj_iters_max']}", flush=True)
    adj.close()
    sol.close()
    assert worst_dot < 1e-9
    assert worst_fd < 1e-7


def test_tangent_linear_check_of_the_512_adjoint():
    """The adjoint case at full size (512^2 cavity, Re 400, 200 steps from
    rest, checkpoint every 20; J = half the squared misfit of the final u, v
    to a target run): on four random directions (two of the lid, two of the
    initial velocity) the FP64 tangent-linear derivative of J equals the
    adjoint gradient's projection to 1e-10 relative (measured 2e-15 ...
    7e-14)."""
    g, reg, prm, x, y = workloads.adjoint_cavity(512)
    nsteps = 200
    run = autograd.CavityRun(g, reg, prm, nsteps, checkpoint_every=20)
    lid_t = workloads.lid_modes(x, y, reg)
    nu_, nv_ = (g.ny - 1, g.nx - 2), (g.ny - 2, g.nx - 1)
    zu, zv = np.zeros(nu_), np.zeros(nv_)
    print("  target run", flush=True)
    uT, vT = run.run(lid_t[2:g.nx + 1], zu, zv, checkpoints=False)
    l0 = np.ones(g.nx - 1)
    uN, vN = run.run(l0, zu, zv)
    gl, gu, gv, st = run.grad(uN - uT, vN - vT)
    assert st["fwd_unconverged"] == 0 and st["adj_unconverged"] == 0
    W = max(g.nx, g.ny) + 2
    rng = np.random.default_rng(512)
    worst = 0.0
    for k in range(4):
        dl = rng.standard_normal(g.nx - 1)
        du, dv = rng.standard_normal(nu_), rng.standard_normal(nv_)
        if k % 2 == 0:
            du[:] = 0.0
            dv[:] = 0.0
        else:
            dl[:] = 0.0
        dbc = np.zeros((4, 2, W))
        dbc[case.NORTH - 1, case.VAR_U - 1, 2:g.nx + 1] = dl
        fu, fv = np.zeros((g.ncol, g.ld)), np.zeros((g.ncol, g.ld))
        fu[run.su] = du
        fv[run.sv] = dv
        t = run.adj.tangent(fu, fv, None, dbc)
        ts = t["stats"]
        assert ts["adj_unconverged"] == 0
        tl = float(np.sum((uN - uT) * t["duN"][run.su]) + np.sum((vN - vT) * t["dvN"][run.sv]))
        ad = float(gl @ dl + np.sum(gu * du) + np.sum(gv * dv))
        rel = abs(tl - ad) / abs(ad)
        worst = max(worst, rel)
        print(f"    direction {k} ({'lid' if k % 2 == 0 else 'u0,v0'}): tangent {tl:.12e} adjoint {ad:.12e} "
              f"rel {rel:.2e}; tangent pass {ts['ms_reverse']:.0f} ms, q-l iterations mean "
              f"{ts['adj_iters_mean']:.1f}, V-cycles mean {ts['ppe_cycles_mean']:.1f}", flush=True)
    run.close()
    assert worst < 1e-10
