"""Boundary profiles on the GPU (wd2_ctx_set_bc_profile) against the oracle.

* A profile of constant values on every side and variable gives the reference
  program's own channel run bit for bit.
* A parabolic inflow (channel) and a sine-series lid (cavity) give the
  oracle's results bit for bit in FP64, and the final ghost fill carries the
  profile exactly.
* The mixed-precision path with a parabolic inflow follows FP64 within 1e-6
  relative L2 per field.
"""
import numpy as np
import pytest

import helpers
from test_profiles import constant_profiles
from wolfd2_amd import api, case
from wolfd2_amd._lib import Wd2Error

pytestmark = pytest.mark.gpu


def _gpu(g, reg, prm, nsteps, precision="fp64"):
    sol = api.Solver(g, reg, prm)
    sol.set_precision(precision)
    z = np.zeros((g.ncol, g.ld))
    sol.set_fields(z, z, z)
    sol.cold_start()
    tab = sol.step(nsteps)
    out = sol.get_fields()
    sol.close()
    return tab, out


def _oracle(oracle, g, reg, prm, nsteps):
    u, v, p = (np.zeros((g.ncol, g.ld)) for _ in range(3))
    oracle.cold_start(reg, g, prm, u, v, p)
    tab = oracle.run(reg, g, prm, u, v, p, nsteps)
    return tab, (u, v, p)


def test_constant_profiles_bitwise_vs_reference_channel():
    fx = helpers.load("run_chan_64x16.npz")
    g, reg, prm = helpers.run_case(fx)
    constant_profiles(reg)
    print("constant profiles, channel 64x16", flush=True)
    tab, (u, v, p) = _gpu(g, reg, prm, int(fx["nts"]))
    ref = fx["table"]
    assert np.array_equal(tab[:, 0], ref[:, 2]) and np.array_equal(tab[:, 1], ref[:, 3])
    S = np.s_[: g.ny + 2, : g.nx + 2]
    assert np.array_equal(u[S], fx["u"]) and np.array_equal(v[S], fx["v"]) and np.array_equal(p[S], fx["p"])


def _parabolic_channel(nx=65, ny=17, solver=5, sorrel=1.0):
    """Channel of aspect 4, Re 500, parabolic inflow of peak 1, mass-conserving
    outlet (with the fully developed outlet the reference's scheme diverges at
    this resolution, with a constant inflow as well)."""
    x, y = case.uniform_nodes(nx, ny, lx=4.0)
    g = case.compute_grid(x, y)
    reg = case.make_regions(nx, ny, inlets=[(1, 1, "west", 1.0, 0.0, 0.0)],
                            outlets=[(1, 1, "east", "mass_cons", 0.0, 0.0, 0.0)])
    s = case.side_positions(x, y, reg, 1, 1, case.WEST, case.VAR_U)
    reg.set_profile(1, 1, case.WEST, case.VAR_U, case.parabolic_profile(s, 1.0))
    prm = case.Params.from_physical(1e-3, 1.0, 1.0 / 500.0, nPpeSolver=solver, sorrel=sorrel)
    return g, reg, prm


@pytest.mark.parametrize("solver", [5, 1])
def test_parabolic_inflow_channel_bitwise_vs_oracle(oracle, solver):
    g, reg, prm = _parabolic_channel(solver=solver)
    print(f"parabolic inflow channel solver={solver}", flush=True)
    tg, (u, v, p) = _gpu(g, reg, prm, 60)
    to, (uo, vo, po) = _oracle(oracle, g, reg, prm, 60)
    assert np.array_equal(tg[:, :2], to[:, :2])
    assert np.array_equal(u, uo) and np.array_equal(v, vo) and np.array_equal(p, po)
    # the inflow face carries the profile on every line of the side
    prof = reg.profiles[(1, 1, case.WEST, case.VAR_U)]
    iw = reg.border(1, 1, case.WEST)
    js, jn = reg.border(1, 1, case.SOUTH), reg.border(1, 1, case.NORTH)
    assert np.array_equal(u[js:jn + 1, iw], prof[js:jn + 1])


def test_sine_lid_cavity_bitwise_vs_oracle(oracle):
    n = 48
    x, y = case.uniform_nodes(n + 1, n + 1)
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(n + 1, n + 1, lid=0.0)
    amps = np.random.default_rng(20261202).normal(0.0, 0.1, 16)
    s = case.side_positions(x, y, reg, 1, 1, case.NORTH, case.VAR_U)
    reg.set_profile(1, 1, case.NORTH, case.VAR_U, case.sine_profile(s, amps))
    prm = case.Params.from_physical(1e-3, 1.0, 1.0 / 400.0, nPpeSolver=5)
    print("sine lid cavity 48^2", flush=True)
    tg, fg = _gpu(g, reg, prm, 40)
    to, fo = _oracle(oracle, g, reg, prm, 40)
    assert np.array_equal(tg[:, :2], to[:, :2])
    for a, b in zip(fg, fo):
        assert np.array_equal(a, b)
    assert np.abs(fg[0]).max() > 0.01   # the lid drives the flow


def test_profile_removed_restores_the_constant():
    fx = helpers.load("run_chan_64x16.npz")
    g, reg, prm = helpers.run_case(fx)
    sol = api.Solver(g, reg, prm)
    sol.set_bc_profile(1, 1, case.WEST, case.VAR_U, np.full(g.ny + 2, 7.0))
    sol.set_bc_profile(1, 1, case.WEST, case.VAR_U, None)
    z = np.zeros((g.ncol, g.ld))
    sol.set_fields(z, z, z)
    sol.cold_start()
    tab = sol.step(20)
    u, v, p = sol.get_fields()
    with pytest.raises(Wd2Error, match="needs 19 values"):
        sol.set_bc_profile(1, 1, case.WEST, case.VAR_U, np.zeros(g.nx + 2))   # west needs ny+2
    sol.close()
    t0, (u0, v0, p0) = _gpu(g, reg, prm, 20)
    assert np.array_equal(tab, t0) and np.array_equal(u, u0) and np.array_equal(p, p0)


def test_mixed_parabolic_channel_follows_fp64():
    g, reg, prm = _parabolic_channel(nx=257, ny=65, sorrel=1.9)
    print("mixed vs fp64 parabolic channel 256x64", flush=True)
    t64, f64 = _gpu(g, reg, prm, 30)
    tmx, fmx = _gpu(g, reg, prm, 30, precision="mixed")
    assert np.array_equal(t64[:, 0], tmx[:, 0])
    S = np.s_[: g.ny + 2, : g.nx + 2]
    for name, a, b in zip("uvp", fmx, f64):
        err = np.linalg.norm(a[S] - b[S]) / np.linalg.norm(b[S])
        print(f"  {name}: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-6, name
