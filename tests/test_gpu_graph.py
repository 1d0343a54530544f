"""hipGraph steps and ensembles (no counterpart in the reference).

* A step replayed as a captured graph equals the host-driven step bit for
  bit, in FP64 and in mixed precision, with and without the multigrid
  solve's extrapolated start (linear and quadratic, also switched during a
  run), including steps whose iteration budget runs
  out and are redone host-driven (the budget follows the last step's counts,
  so the first steps of a cold start rise past it).
* An ensemble of cavities of uneven sizes and Reynolds numbers, advanced
  together (graph mode and host-driven), gives every member the bits it has
  when it runs alone.
"""
import numpy as np
import pytest

from wolfd2_amd import api, case, workloads

pytestmark = pytest.mark.gpu


def _solver(g, reg, prm, precision, graph, fields=None, extrap=0):
    sol = api.Solver(g, reg, prm)
    sol.set_precision(precision)
    sol.set_graph(graph)
    sol.set_mg_extrapolation(extrap)
    if fields is None:
        z = np.zeros((g.ncol, g.ld))
        sol.set_fields(z, z, z)
        sol.cold_start()
    else:
        sol.set_fields(*fields)
    return sol


@pytest.mark.parametrize("precision,extrap", [("fp64", 0), ("mixed", 0), ("fp64", 2), ("mixed", 1), ("mixed", 2)])
def test_graph_steps_equal_host_driven_steps(precision, extrap):
    g, reg, prm = workloads.cavity(128, 1000.0, 1e-3, nPpeSolver=7, msorit=60, sortol=1e-10)
    out = {}
    for graph in (False, True):
        sol = _solver(g, reg, prm, precision, graph, extrap=extrap)
        tab = sol.step(30)
        out[graph] = (tab, sol.get_fields(), sol.graph_stats())
        sol.close()
    (t0, f0, _), (t1, f1, st) = out[False], out[True]
    print(f"  {precision}, extrapolation order {extrap}: graph steps {st}, ql {t1[:6, 0].tolist()} "
          f"cycles {t1[:6, 1].tolist()}", flush=True)
    assert st["replayed"] > 0 and st["replayed"] + st["redone"] == 30
    assert st["redone"] >= 1, "a cold start's rising counts should exceed the first budgets"
    assert np.array_equal(t0, t1)
    for a, b in zip(f0, f1):
        assert np.array_equal(a, b)


def test_graph_steps_channel_with_profile_and_outlet():
    g, reg, prm, x, y = workloads.channel(256, 64, outlet="mass_cons", dt=5e-4, nPpeSolver=7, msorit=60,
                                          sortol=1e-10)
    f0 = workloads.poiseuille_noise(g, reg, x, y)
    out = {}
    for graph in (False, True):
        sol = _solver(g, reg, prm, "fp64", graph, fields=f0)
        out[graph] = (sol.step(20), sol.get_fields())
        sol.close()
    assert np.array_equal(out[False][0], out[True][0])
    for a, b in zip(out[False][1], out[True][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_members_equal_their_single_runs(graph):
    specs = workloads.ensemble_specs(6, seed=7, sides=(64, 96, 128, 192))
    precisions = ["mixed", "fp64"] * 3
    cases = [workloads.cavity(n, re, 5e-4, nPpeSolver=7, msorit=60, sortol=1e-10) for n, re in specs]
    print(f"  ensemble members (side, Re): {[(n, round(r)) for n, r in specs]}", flush=True)
    alone = []
    for (g, reg, prm), prec in zip(cases, precisions):
        sol = _solver(g, reg, prm, prec, graph)
        alone.append((sol.step(12), sol.get_fields()))
        sol.close()
    members = [_solver(g, reg, prm, prec, graph) for (g, reg, prm), prec in zip(cases, precisions)]
    ens = api.Ensemble(members)
    tabs = ens.step(12)
    ens.close()
    for i, m in enumerate(members):
        assert np.array_equal(tabs[i], alone[i][0]), i
        for a, b in zip(m.get_fields(), alone[i][1]):
            assert np.array_equal(a, b), i
        m.close()


def test_graph_cache_follows_precision_params_and_profiles():
    """Cached step graphs bake in the precision, the step parameters and the
    boundary tables and the extrapolated start's order and history slot:
    after each change the graph steps still equal host-driven steps bit for
    bit, and they replay (a stale graph would fail its stopping test and be
    redone, or give other fields)."""
    g, reg, prm = workloads.cavity(96, 400.0, 5e-4, nPpeSolver=7, msorit=100, sortol=1e-11)
    z = np.zeros((g.ncol, g.ld))
    x, y = case.uniform_nodes(g.nx, g.ny)
    lid2 = workloads.lid_modes(x, y, reg, seed=3)
    prm2 = case.Params.from_physical(3e-4, 1.0, 1.0 / 700.0, nPpeSolver=7, msorit=100, sortol=1e-11)

    def run(graph):
        sol = api.Solver(g, reg, prm)
        sol.set_graph(graph)
        sol.set_fields(z, z, z)
        tabs = [sol.step(6)]
        sol.set_precision("mixed")
        tabs.append(sol.step(6))
        sol.set_params(prm2)
        tabs.append(sol.step(6))
        sol.set_mg_extrapolation(2)
        sol.set_bc_profile(1, 1, case.NORTH, case.VAR_U, lid2)
        tabs.append(sol.step(6))
        sol.set_mg_extrapolation(1)
        sol.set_precision("fp64")
        tabs.append(sol.step(6))
        out = sol.get_fields(), np.concatenate(tabs), sol.graph_stats()
        sol.close()
        return out

    (fe, te, _), (fg, tg, st) = run(False), run(True)
    print(f"  graph stats {st}", flush=True)
    assert np.array_equal(te, tg)
    for a, b in zip(fe, fg):
        assert np.array_equal(a, b)
    # each of the five phases redoes at most its first two steps
    assert st["replayed"] >= 20


@pytest.mark.parametrize("precision,order", [("fp64", 1), ("mixed", 1), ("fp64", 2), ("mixed", 2)])
def test_mg_extrapolated_start_converges_to_the_same_solution(precision, order):
    """With the extrapolated start (order 1: 2 p^n - p^(n-1), order 2:
    3 p^n - 3 p^(n-1) + p^(n-2)) every step's multigrid solve still meets the
    divergence tolerance, the cavity's fields stay close to the plain start's
    after 30 steps (both solves stop at a divergence of 1e-10): u and v within
    1e-7 of the velocity's norm, p within 1e-6 of its own (relative L2); and
    the V-cycle count does not rise."""
    g, reg, prm = workloads.cavity(128, 1000.0, 1e-3, nPpeSolver=7, msorit=60, sortol=1e-10)
    out = {}
    for extrap in (0, order):
        sol = _solver(g, reg, prm, precision, False, extrap=extrap)
        if extrap == 2:
            with pytest.raises(api.Wd2Error):
                sol.set_mg_extrapolation(3)
        tab = sol.step(30)
        out[extrap] = (tab, sol.get_fields())
        sol.close()
    (t0, f0), (t1, f1) = out[0], out[order]
    print(f"  {precision}, order {order}: V-cycles per step plain {t0[:, 1].mean():.2f}, extrapolated "
          f"{t1[:, 1].mean():.2f}", flush=True)
    assert np.all(t1[:, 1] < 60), "a solve hit the cycle cap"
    assert t1[:, 1].sum() <= t0[:, 1].sum()
    S = np.s_[: g.ny + 2, : g.nx + 2]
    # u and v against the norm of the velocity (v is small in the first
    # steps from rest), p against its own
    speed = np.sqrt(np.linalg.norm(f0[0][S]) ** 2 + np.linalg.norm(f0[1][S]) ** 2)
    errs = {}
    for name, a, b in zip("uvp", f1, f0):
        errs[name] = np.linalg.norm(a[S] - b[S]) / (speed if name != "p" else np.linalg.norm(b[S]))
        print(f"    {name}: extrapolated vs plain {errs[name]:.2e}", flush=True)
    assert errs["u"] < 1e-7 and errs["v"] < 1e-7 and errs["p"] < 1e-6


def test_mg_extrapolated_start_applies_two_cycles_at_least():
    """A solve from the extrapolated start applies at least two V-cycles even
    where one would meet the tolerance (with one, the error in the modes the
    divergence test barely sees grew over thousands of steps of a 1024^2
    cavity).  400 steps of a 128^2 cavity with the quadratic start: every
    extrapolated solve (steps 2..) applies two cycles or more (about a fifth
    stop at two), u and v stay within 1e-7 of the plain start's against the
    velocity's norm (measured 4.9e-8, 3.3e-8), and p within 2e-5 of the
    plain start's norm after the mean of the difference is taken out
    (measured 9.3e-6).  Both runs meet the divergence tolerance of 1e-10 at
    every step; the smooth pressure modes that barely change the divergence
    are left free within it, and each step carries its differences in them
    over to the next through the ghost values, so p differs more than the
    velocity does."""
    g, reg, prm = workloads.cavity(128, 1000.0, 1e-3, nPpeSolver=7, msorit=60, sortol=1e-10)
    out = {}
    for extrap in (0, 2):
        sol = _solver(g, reg, prm, "mixed", False, extrap=extrap)
        tab = sol.step(400)
        out[extrap] = (tab, sol.get_fields())
        sol.close()
    (t0, f0), (t2, f2) = out[0], out[2]
    counts = np.bincount(t2[1:, 1].astype(int))
    print(f"  V-cycles per step from the quadratic start (steps 2..): {dict(enumerate(counts.tolist()))}; "
          f"plain start mean {t0[:, 1].mean():.2f}", flush=True)
    assert np.all(t2[1:, 1] >= 2)
    assert np.all(t2[:, 1] < 60), "a solve hit the cycle cap"
    S = np.s_[: g.ny + 2, : g.nx + 2]
    speed = np.sqrt(np.linalg.norm(f0[0][S]) ** 2 + np.linalg.norm(f0[1][S]) ** 2)
    for name, a, b in zip("uv", f2[:2], f0[:2]):
        e = np.linalg.norm(a[S] - b[S]) / speed
        print(f"    {name}: quadratic vs plain start after 400 steps {e:.2e}", flush=True)
        assert e < 1e-7
    P = np.s_[2: g.ny + 1, 2: g.nx + 1]   # the pressure's cells
    d = f2[2][P] - f0[2][P]
    nb = np.linalg.norm(f0[2][P])
    raw, e = np.linalg.norm(d) / nb, np.linalg.norm(d - d.mean()) / nb
    print(f"    p: quadratic vs plain start after 400 steps {e:.2e} without the mean ({raw:.2e} with it)", flush=True)
    assert e < 2e-5
