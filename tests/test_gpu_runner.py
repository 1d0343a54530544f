"""The deck runner (python -m wolfd2_amd.run): the reference's own case files
in, the reference's PrintDiff table and restart file out, with the whole run on
the GPU.  Each fixture deck must reproduce the reference program's table and
final fields bit for bit, and a run continued from its restart file must equal
the uninterrupted run."""
import io
import os

import numpy as np
import pytest

import helpers
import refdecks
from wolfd2_amd import deck
from wolfd2_amd.run import run_case

pytestmark = pytest.mark.gpu


def write_case(rd, fx, text=None):
    refdecks.write_grid(os.path.join(rd, "grid.xyz"), fx["x"], fx["y"])
    refdecks.write_fluid_table(os.path.join(rd, "fluid.dat"), "fl01", float(fx["rho"]), float(fx["mu"]))
    with open(os.path.join(rd, "input.dat"), "w") as f:
        f.write(str(fx["deck"]) if text is None else text)


# every deck except the two 1000-step / line-SOR decks whose whole runs
# test_gpu_parity.py already checks through the same context (they are slow)
RUNNER_CASES = [n for n in helpers.run_fixture_names() if n not in ("run_c1_sor_64.npz", "run_c1_lsor_32.npz")]


@pytest.mark.parametrize("name", RUNNER_CASES)
def test_runner_reproduces_the_reference_program(name, tmp_path):
    fx = helpers.load(name)
    write_case(str(tmp_path), fx)
    print(f"runner {name} steps={int(fx['nts'])}", flush=True)
    buf = io.StringIO()
    r = run_case(str(tmp_path), out=buf)
    ref = fx["table"]
    tab = refdecks.parse_print_diff(buf.getvalue())
    fin = np.all(np.isfinite(ref), axis=1)
    n = min(len(tab), len(ref))
    assert n > 0
    assert np.array_equal(tab[:n][fin[:n]], ref[:n][fin[:n]]), "PrintDiff table differs"
    if np.all(np.isfinite(fx["u"])):
        rst = deck.read_restart(str(tmp_path / "out.rst"))
        assert rst["k"] == int(fx["k"]) and rst["dtime"] == float(fx["dtime"])
        for f in ("u", "v", "p"):
            assert np.array_equal(rst[f], fx[f]), f


def test_runner_restart_continues_the_run(tmp_path):
    fx = helpers.load("run_c5_re1000_128.npz")
    text = str(fx["deck"])
    full, first, second = (tmp_path / d for d in ("full", "first", "second"))
    for d in (full, first, second):
        d.mkdir()
    write_case(str(full), fx, text.replace("n_time_steps 50", "n_time_steps 30"))
    write_case(str(first), fx, text.replace("n_time_steps 50", "n_time_steps 12"))
    second_text = text.replace("n_time_steps 50", "n_time_steps 18").replace(
        "write_restart out.rst", "restart in.rst\nwrite_restart out.rst")
    write_case(str(second), fx, second_text)
    b_full, b1, b2 = io.StringIO(), io.StringIO(), io.StringIO()
    r_full = run_case(str(full), out=b_full)
    run_case(str(first), out=b1)
    os.replace(first / "out.rst", second / "in.rst")
    r2 = run_case(str(second), out=b2)
    t_full = refdecks.parse_print_diff(b_full.getvalue())
    t_split = np.vstack([refdecks.parse_print_diff(b1.getvalue()), refdecks.parse_print_diff(b2.getvalue())])
    assert np.array_equal(t_full, t_split)
    a, b = deck.read_restart(str(full / "out.rst")), deck.read_restart(str(second / "out.rst"))
    assert (a["k"], a["dtime"]) == (b["k"], b["dtime"]) == (30, r_full["dtime"])
    for f in ("u", "v", "p"):
        assert np.array_equal(a[f], b[f])


def test_runner_mixed_precision(tmp_path):
    fx = helpers.load("run_c5_re1000_128.npz")
    write_case(str(tmp_path), fx)
    r64 = run_case(str(tmp_path), out=io.StringIO())
    rmx = run_case(str(tmp_path), precision="mixed", out=io.StringIO())
    for f in ("u", "v", "p"):
        a, b = rmx[f], r64[f]
        assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b), f


def test_runner_bc_profile_deck_bitwise_vs_oracle(oracle, tmp_path):
    """A channel deck with a parabolic inflow (bc_profile) runs on the GPU and
    equals the oracle on the case the same deck builds."""
    from wolfd2_amd import case
    fx = helpers.load("run_chan_64x16.npz")
    text = str(fx["deck"]).replace("n_time_steps 200", "n_time_steps 60").replace(
        "outlet 1 1 east mass_cons", "outlet 1 1 east mass_cons\nbc_profile 1 1 west u parabolic 1.0")
    write_case(str(tmp_path), fx, text)
    print("runner bc_profile channel 64x16 steps=60", flush=True)
    run_case(str(tmp_path), out=io.StringIO())
    rst = deck.read_restart(str(tmp_path / "out.rst"))
    d = deck.read_input(str(tmp_path / "input.dat"))
    x, y = deck.read_grid(str(tmp_path / "grid.xyz"))
    ny, nx = x.shape
    reg = deck.build_case(d, float(fx["rho"]), nx, ny, x=x, y=y, case_dir=str(tmp_path))
    assert (1, 1, case.WEST, case.VAR_U) in reg.profiles
    g, _, prm = helpers.run_case(fx)
    u, v, p = (np.zeros((g.ncol, g.ld)) for _ in range(3))
    oracle.cold_start(reg, g, prm, u, v, p)
    oracle.run(reg, g, prm, u, v, p, 60)
    S = np.s_[: ny + 2, : nx + 2]
    for f, a in zip("uvp", (u, v, p)):
        assert np.array_equal(rst[f], a[S]), f


def test_runner_thermal_reproduces_the_reference_program(tmp_path):
    """A heated cavity deck (energy equation, buoyancy, filters, momentum/
    energy iterations, heat generation) on the device-resident context: the
    PrintDiff lines (four norms) and the restart fields u, v, p, t equal the
    reference program's."""
    fx = helpers.load("thermrun_therm_cavity_32.npz")
    refdecks.write_grid(str(tmp_path / "grid.xyz"), fx["x"], fx["y"])
    refdecks.write_fluid_table(str(tmp_path / "fluid.dat"), "fl01", float(fx["rho"]), float(fx["mu"]),
                               float(fx["gas_constant"]))
    (tmp_path / "input.dat").write_text(str(fx["deck"]))
    print(f"runner thermal cavity steps={int(fx['nts'])}", flush=True)
    buf = io.StringIO()
    run_case(str(tmp_path), out=buf)
    lines = [ln for ln in buf.getvalue().splitlines() if refdecks._DIFF_RE.match(ln)]
    assert lines == str(fx["diff_lines"]).split("\n"), "PrintDiff lines differ"
    rst = deck.read_restart(str(tmp_path / "out.rst"))
    assert rst["k"] == int(fx["k"]) and rst["dtime"] == float(fx["dtime"])
    for f in ("u", "v", "p", "t"):
        assert np.array_equal(rst[f], fx[f]), f
