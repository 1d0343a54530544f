"""The product's reader of the reference's input files (wolfd2_amd.deck).

Pinned by the fixtures' decks (the same regions and parameters as the test
helpers build, which the whole-run parity tests check against the reference),
by the compiled reference program's own PrintDiff lines (byte for byte), and by
closed forms (fluid table interpolation, restart round trip).
"""
import io
import os
import subprocess
import tempfile

import numpy as np
import pytest

import helpers
import refdecks
from wolfd2_amd import case, deck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "std", "wolfd2")


def write_case(rd, fx):
    refdecks.write_grid(os.path.join(rd, "grid.xyz"), fx["x"], fx["y"])
    refdecks.write_fluid_table(os.path.join(rd, "fluid.dat"), "fl01", float(fx["rho"]), float(fx["mu"]))
    with open(os.path.join(rd, "input.dat"), "w") as f:
        f.write(str(fx["deck"]))


@pytest.mark.parametrize("name", helpers.run_fixture_names())
def test_deck_gives_the_fixture_case(name, tmp_path):
    fx = helpers.load(name)
    write_case(str(tmp_path), fx)
    d = deck.read_input(str(tmp_path / "input.dat"))
    s = d.settings
    x, y = deck.read_grid(str(tmp_path / s["grid_file"]))
    assert np.array_equal(x, fx["x"]) and np.array_equal(y, fx["y"])
    fl = deck.read_fluid(str(tmp_path / s["flPropFn"]), s["flName"], s["tref"])
    assert fl[1] == float(fx["rho"]) and fl[3] == float(fx["mu"])
    ny, nx = x.shape
    reg = deck.build_case(d, float(fl[1]), nx, ny)
    prm = case.Params.from_physical(s["dk"], float(fl[1]), float(fl[3]), mqiter=s["mqiter"], qtol=s["qtol"],
                                    nPpeSolver=s["nPpeSolver"], msorit=s["msorit"], sortol=s["sortol"],
                                    sorrel=s["sorrel"], cartesian=s["cartesian"])
    g0, reg0, prm0 = helpers.run_case(fx)
    assert prm == prm0
    for f in ("nreg", "nRegBrd", "nRegType", "nMomBdTp", "dBCVal", "dPRporos"):
        assert np.array_equal(getattr(reg, f), getattr(reg0, f)), f
    assert s["nts"] == int(fx["nts"])


def test_fluid_table_interpolation(tmp_path):
    p = tmp_path / "f.dat"
    p.write_text("fluid a\n200.0 1.0 1.0 2.0 0.1 0.7\n400.0 2.0 1.0 4.0 0.1 0.7\nend\n"
                 "fluid b\n100.0 9.0 1.0 9.0 0.1 0.7\n900.0 9.0 1.0 9.0 0.1 0.7\nend\n")
    fl = deck.read_fluid(str(p), "a", 300.0)
    assert fl[0] == 300.0 and fl[1] == 1.5 and fl[3] == 3.0
    assert deck.read_fluid(str(p), "b", 283.57)[1] == 9.0
    with pytest.raises(deck.DeckError):
        deck.read_fluid(str(p), "a", 150.0)


def test_restart_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    f = {k: rng.standard_normal((7, 9)) for k in deck.RESTART_FIELDS}
    deck.write_restart(str(tmp_path / "r.rst"), 42, 0.125, 7, 5, f)
    r = deck.read_restart(str(tmp_path / "r.rst"))
    assert (r["k"], r["dtime"], r["nx"], r["ny"]) == (42, 0.125, 7, 5)
    for k in deck.RESTART_FIELDS:
        assert np.array_equal(r[k], f[k])


def test_refused_features(tmp_path):
    p = tmp_path / "input.dat"
    p.write_text("section input_parameters\ngrid_file g.xyz\nsmall_scale\nend section\n")
    with pytest.raises(deck.DeckError, match="outside the GPU hot path"):
        deck.read_input(str(p))
    p.write_text("section input_parameters\ngrid_file g.xyz\nend section\nsection boundary_conditions\n"
                 "number_of_regions 1 1\nporous_region 1 1 kaviany 0.4 0.001 0.55\nend section\n")
    with pytest.raises(deck.DeckError, match="porous_region"):
        deck.read_input(str(p))


@pytest.mark.parametrize("x,s", [(0.0, " 0.000000E+00"), (1.0, " 0.100000E+01"), (-2.5e-7, "-0.250000E-06"),
                                 (123456.789, " 0.123457E+06"), (9.9999996e-3, " 0.100000E-01")])
def test_fortran_e_edit(x, s):
    assert deck.fortran_e(x, 13, 6) == s


@pytest.mark.skipif(not os.path.exists(REF_EXE), reason="oracle/_ref not built (needs the reference sources)")
def test_print_diff_lines_match_the_reference_program():
    """The reference program's PrintDiff lines, reformatted from their parsed
    values by deck.print_diff_row, come out byte for byte the same."""
    nx = ny = 17
    x, y = refdecks.uniform_grid(nx, ny)
    d = refdecks.Deck(nts=12, dt=1e-3, mu=0.01, ppe_solver="rb_sor")
    with tempfile.TemporaryDirectory() as rd:
        d.write(rd, x, y)
        out = subprocess.run([REF_EXE], cwd=rd, capture_output=True, text=True, timeout=60, check=True).stdout
    lines = [ln for ln in out.splitlines() if refdecks._DIFF_RE.match(ln)]
    assert len(lines) == 12
    for ln in lines:
        r = refdecks.parse_print_diff(ln)[0]
        mine = deck.print_diff_row(int(r[0]), r[1], int(r[2]), int(r[3]), r[4:7])
        assert mine.rstrip() == ln.rstrip(), (mine, ln)


THERMAL_DECK = dict(nts=12, dt=1e-3, mu=0.01, ppe_solver="rb_sor", nreg=(2, 1), i_borders=(9,), gas_constant=287.0,
                    extra_input=["thermal_energy", "buoyancy", "max_me_iter 2", "filter_t 30.0",
                                 "ref_temperature 300.0", "max_ref_temperature 320.0", "print_bann_freq 5"],
                    bc_lines=["wall 1 1 west temper 320.0", "fixed_temperature_wall 2 1 east 310.0",
                              "heat_flux_wall 1 1 south", "heat_generation_region 2 1 25.0",
                              "wall 1 1 north tangent_vel 0.5"])


@pytest.mark.skipif(not os.path.exists(REF_EXE), reason="oracle/_ref not built (needs the reference sources)")
def test_print_diff_thermal_lines_match_the_reference_program():
    """With the energy equation the reference prints a 78-column rule, four
    a12 headings and rows of four e12.5 norms; the banner and rows of
    deck.PRINTDIFF_*_T and deck.print_diff_row(thermal=True) match them."""
    nx = ny = 17
    x, y = refdecks.uniform_grid(nx, ny)
    d = refdecks.Deck(**THERMAL_DECK)
    with tempfile.TemporaryDirectory() as rd:
        d.write(rd, x, y)
        out = subprocess.run([REF_EXE], cwd=rd, capture_output=True, text=True, timeout=60, check=True).stdout
    lines = out.splitlines()
    assert deck.PRINTDIFF_HEAD_T.rstrip() in [ln.rstrip() for ln in lines]
    assert sum(ln == deck.PRINTDIFF_RULE_T for ln in lines) == 2 * 3   # banners before steps 1, 6, 11
    rows = [ln for ln in lines if refdecks._DIFF_RE.match(ln)]
    assert len(rows) == 12
    for ln in rows:
        w = ln.split()
        k, t, ql = int(w[0]), float(w[1]), int(w[2].rstrip("*"))
        mine = deck.print_diff_row(k, t, ql, int(w[3]), [float(v) for v in w[4:8]], thermal=True)
        assert mine == ln, (mine, ln)


def test_thermal_deck_tables(tmp_path):
    """The thermal tables the deck builds: defaults (adiabatic outer walls),
    wall temperatures scaled by (T - tref)/(tmax - tref), a fixed-temperature
    wall, heat generation made non-dimensional by NDHGenCoef."""
    nx = ny = 17
    x, y = refdecks.uniform_grid(nx, ny)
    dk = refdecks.Deck(**THERMAL_DECK)
    dk.write(str(tmp_path), x, y)
    d = deck.read_input(str(tmp_path / "input.dat"))
    s = d.settings
    assert (s["nthermen"], s["neqstate"], s["nmeiter"], s["nfiltt"], s["fpt"]) == (1, 1, 2, 1, 30.0)
    assert (s["tref"], s["tmax"]) == (300.0, 320.0)
    assert deck.read_gas_constant(str(tmp_path / "fluid.dat"), "fl01") == 287.0
    reg = deck.build_case(d, 1.0, nx, ny)
    th = deck.build_thermal(d, reg, 0.025, 70.0)
    W, E, S, N = case.WEST - 1, case.EAST - 1, case.SOUTH - 1, case.NORTH - 1
    assert th.nTemBdTp[0, 0, W] == case.BT_TEMPER and reg.dBCVal[0, 0, W, 3] == (320.0 - 300.0) / 20.0
    assert th.nTemBdTp[1, 0, E] == case.BT_TEMPER and reg.dBCVal[1, 0, E, 3] == (310.0 - 300.0) / 20.0
    assert th.nTemBdTp[0, 0, S] == case.BT_HTFLUX and th.nTemBdTp[1, 0, N] == case.BT_HTFLUX
    assert th.nTemBdTp[0, 0, E] == case.BT_INTERN and th.nTemBdTp[1, 0, W] == case.BT_INTERN
    assert th.nTRgType[1, 0] == case.RT_HEATGN and th.nTRgType[0, 0] == case.RT_NOSRCE
    assert th.dHGSTval[1, 0] == (25.0 * (1.0 * 1.0)) / (70.0 * 0.025 * (320.0 - 300.0))
