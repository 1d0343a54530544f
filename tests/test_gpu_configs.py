"""The measured workloads at their own sizes, against the oracle and between
the two arithmetic paths.

* The 1024^2 cavity at the bench's exact settings (Re 1000, dt 2e-4, rb_sor,
  the reference's tolerances: the Poisson solve runs to its 2000-sweep cap,
  and the solve after a capped one starts with the long batches of launches):
  cold start and 2 steps, FP64 bit for bit against the oracle.
* The 2048 x 512 channel (Lx = 4, Re 500, dt 1e-4, parabolic inflow of peak
  1, Poiseuille plus N(0, 1e-3^2) noise given through a restart file): FP64
  bit for bit against the oracle for 3 steps with the zero-gradient outlet
  (OUTLT1, bound_cond.f:1656-1874) and with the mass-conserving one (OUTLT2).
  With the zero-gradient outlet the reference's own scheme diverges at this
  size (the oracle's |p1-pn| grows about 2.4x a step, max|u| reaches 38 at
  step 8 and NaN at step 9), so the 100-step comparison of the mixed path
  with FP64 (within 1e-6 relative L2 per field) runs the mass-conserving
  outlet, which stays at max|u| = 1.001 over 30 oracle steps.

Bars: bit-exact (np.array_equal) for FP64; relative L2 <= 1e-6 per field for
mixed precision.
"""
import numpy as np
import pytest

from wolfd2_amd import api, workloads
from wolfd2_amd._lib import lib

pytestmark = pytest.mark.gpu

TOL_L2 = 1e-6


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def test_cavity_1024_bench_settings_bitwise_vs_oracle(oracle):
    g, reg, prm = workloads.cavity(1024, 1000.0, 2e-4, nPpeSolver=5)
    print("cavity 1024^2, bench settings, cold start + 2 steps (gpu)", flush=True)
    sol = api.Solver(g, reg, prm)
    z = np.zeros((g.ncol, g.ld))
    sol.set_fields(z, z, z)
    n0 = sol.cold_start()
    tg = sol.step(2)
    fg = sol.get_fields()
    sol.close()
    print(f"  gpu: cold start {n0} sweeps, steps {tg[:, :2].tolist()}; oracle...", flush=True)
    u, v, p = (np.zeros((g.ncol, g.ld)) for _ in range(3))
    m0 = oracle.cold_start(reg, g, prm, u, v, p)
    to = oracle.run(reg, g, prm, u, v, p, 2)
    assert n0 == m0
    assert np.all(to[:, 1] == prm.msorit), "the bench's solves run to the sweep cap"
    assert np.array_equal(tg[:, :2], to[:, :2])
    assert np.array_equal(tg[:, 2:], to[:, 2:])
    for a, b in zip(fg, (u, v, p)):
        assert np.array_equal(a, b)


def _channel(outlet="fully_dev"):
    g, reg, prm, x, y = workloads.channel(2048, 512, outlet=outlet, nPpeSolver=5)
    u0, v0, p0 = workloads.through_restart(g, *workloads.poiseuille_noise(g, reg, x, y))
    return g, reg, prm, (u0, v0, p0)


def _gpu(g, reg, prm, fields, nsteps, precision="fp64"):
    sol = api.Solver(g, reg, prm)
    sol.set_precision(precision)
    sol.set_fields(*fields)
    tab = sol.step(nsteps)
    out = sol.get_fields()
    sol.close()
    return tab, out


@pytest.mark.parametrize("outlet", ["fully_dev", "mass_cons"])
def test_channel_2048x512_bitwise_vs_oracle(oracle, outlet):
    g, reg, prm, f0 = _channel(outlet)
    print(f"channel 2048x512 {outlet}, 3 steps from Poiseuille+noise (gpu)", flush=True)
    tg, fg = _gpu(g, reg, prm, f0, 3)
    print(f"  gpu steps {tg.tolist()}; oracle...", flush=True)
    u, v, p = (a.copy() for a in f0)
    to = oracle.run(reg, g, prm, u, v, p, 3)
    assert np.array_equal(tg[:, :2], to[:, :2])
    assert np.array_equal(tg[:, 2:], to[:, 2:])
    for a, b in zip(fg, (u, v, p)):
        assert np.array_equal(a, b)
    assert np.all(np.isfinite(fg[0]))


def test_channel_2048x512_mixed_follows_fp64():
    g, reg, prm, f0 = _channel("mass_cons")
    steps = 100
    print(f"channel 2048x512 mass_cons, mixed vs fp64, {steps} steps", flush=True)
    t64, f64 = _gpu(g, reg, prm, f0, steps)
    tmx, fmx = _gpu(g, reg, prm, f0, steps, precision="mixed")
    assert np.all(np.isfinite(f64[0])) and np.all(np.isfinite(fmx[0]))
    assert np.array_equal(t64[:, 0], tmx[:, 0]), "q-l iteration counts differ"
    S = np.s_[: g.ny + 2, : g.nx + 2]
    # the velocity components against the velocity scale ||(u, v)|| (v is
    # the inflow's noise, three orders below u, so its own norm would measure
    # FP32 rounding of u-sized terms against 1e-3-sized values); p against ||p||
    speed = np.sqrt(np.linalg.norm(f64[0][S]) ** 2 + np.linalg.norm(f64[1][S]) ** 2)
    for name, a, b in zip("uvp", fmx, f64):
        d = np.linalg.norm(a[S] - b[S])
        e = d / (speed if name in "uv" else np.linalg.norm(b[S]))
        print(f"  {name}: rel L2 {e:.3e} (against its own norm {rel_l2(a[S], b[S]):.3e}), "
              f"max {np.abs(a[S] - b[S]).max():.3e}", flush=True)
        assert e <= TOL_L2, name
    print(f"  sequential chain fallbacks so far: {lib().wd2_trid_fallbacks()}", flush=True)
