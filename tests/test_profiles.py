"""Boundary profiles: one value per grid line along a region side in place of
the reference's constant boundary value (dBCVal), in the oracle and the deck.

The reference has no profiles, so nothing it wrote covers them directly.
They are pinned by:

* the constant case: a profile of constant values on every side and variable
  reproduces the reference program's own results (kernel goldens of every
  boundary type, a whole channel run) bit for bit;
* locality: changing the value of grid line k changes the boundary fill on
  line k of that side, and elsewhere only at the corners where a later side
  reads it;
* closed forms of where each line's value applies (cell centres or nodes) and
  of the inflow a parabolic profile carries.
"""
import numpy as np
import pytest

import helpers
from wolfd2_amd import case, deck

VARS = (case.VAR_U, case.VAR_V, case.VAR_P)
SIDES = (case.WEST, case.EAST, case.SOUTH, case.NORTH)


def constant_profiles(reg):
    """Every side of every region gets profiles equal to its constants."""
    ni, nj = int(reg.nreg[0]), int(reg.nreg[1])
    for jr in range(1, nj + 1):
        for ir in range(1, ni + 1):
            for side in SIDES:
                for var in VARS:
                    c = reg.dBCVal[ir - 1, jr - 1, side - 1, var - 1]
                    reg.set_profile(ir, jr, side, var, np.full(reg.profile_length(side), c))


@pytest.mark.parametrize("name", helpers.kernel_fixture_names())
def test_constant_profiles_give_the_reference_boundary_fills(oracle, name):
    fx = helpers.load(name)
    nx, ny = int(fx["nx"]), int(fx["ny"])
    g = case.compute_grid(fx["x"], fx["y"])
    reg = helpers.regions_from(fx)
    # distinct non-zero constants, so that a lookup of the wrong side or
    # variable shows; the expected fills come from the oracle without profiles,
    # which the reference's goldens pin
    rng = np.random.default_rng(5)
    reg.dBCVal = np.asfortranarray(reg.dBCVal + rng.uniform(0.5, 1.5, reg.dBCVal.shape))
    ld, nc = g.ld, g.ncol
    E = lambda a: helpers.embed(a, ld, nc)  # noqa: E731
    base = []
    for prof in (False, True):
        if prof:
            constant_profiles(reg)
        u, v = E(fx["vbc_u_in"]), E(fx["vbc_v_in"])
        assert oracle.vel_bound_cond(reg, ld, u, v) == 0
        u2, v2 = E(fx["vob_u_in"]), E(fx["vob_v_in"])
        oracle.vel_outflow_bcs(reg, ld, u2, v2)
        p = E(fx["pbc_p_in"])
        oracle.pres_bound_cond(reg, ld, p)
        out = (u, v, u2, v2, p)
        if not prof:
            base = out
        else:
            for a, b in zip(out, base):
                assert np.array_equal(a, b)
    # and the unshifted constants still give the reference's own fills
    reg = helpers.regions_from(fx)
    constant_profiles(reg)
    u, v = E(fx["vbc_u_in"]), E(fx["vbc_v_in"])
    oracle.vel_bound_cond(reg, ld, u, v)
    S = np.s_[: ny + 2, : nx + 2]
    assert np.array_equal(u[S], fx["vbc_u_out"]) and np.array_equal(v[S], fx["vbc_v_out"])
    p = E(fx["pbc_p_in"])
    oracle.pres_bound_cond(reg, ld, p)
    assert np.array_equal(p[S], fx["pbc_p_out"])


def test_constant_profiles_reproduce_the_reference_channel_run(oracle):
    fx = helpers.load("run_chan_64x16.npz")
    g, reg, prm = helpers.run_case(fx)
    constant_profiles(reg)
    u, v, p = (np.zeros((g.ncol, g.ld)) for _ in range(3))
    oracle.cold_start(reg, g, prm, u, v, p)
    table = oracle.run(reg, g, prm, u, v, p, int(fx["nts"]))
    ref = fx["table"]
    assert np.array_equal(table[:, 0], ref[:, 2]) and np.array_equal(table[:, 1], ref[:, 3])
    S = np.s_[: g.ny + 2, : g.nx + 2]
    assert np.array_equal(u[S], fx["u"]) and np.array_equal(v[S], fx["v"]) and np.array_equal(p[S], fx["p"])


def _locality_layout(nx=13, ny=9):
    """One region: inlet west, moving walls south and north, fully developed
    outlet east (the mass-conserving outlet reads no boundary velocity)."""
    return case.make_regions(nx, ny, walls=[(1, 1, "south", "no_slip", 0.3, 0.0), (1, 1, "north", "no_slip", 1.0, 0.0)],
                             inlets=[(1, 1, "west", 1.0, 0.2, 0.0)],
                             outlets=[(1, 1, "east", "fully_dev", 0.0, 0.0, 0.4)])


@pytest.mark.parametrize("side", SIDES)
@pytest.mark.parametrize("var", VARS)
def test_profile_line_changes_only_its_grid_line(oracle, side, var):
    nx, ny = 13, 9
    x, y = case.uniform_nodes(nx, ny)
    g = case.compute_grid(x, y)
    reg = _locality_layout(nx, ny)
    rng = np.random.default_rng(11)
    u0, v0, p0 = (rng.standard_normal((g.ncol, g.ld)) for _ in range(3))

    def fill(r):
        u, v, p = u0.copy(), v0.copy(), p0.copy()
        assert oracle.vel_bound_cond(r, g.ld, u, v) == 0
        assert oracle.vel_outflow_bcs(r, g.ld, u, v) == 0
        assert oracle.pres_bound_cond(r, g.ld, p) == 0
        return u, v, p

    constant_profiles(reg)
    base = fill(reg)
    n = reg.profile_length(side)
    moved = set()
    for k in range(n):
        vals = np.full(n, reg.dBCVal[0, 0, side - 1, var - 1])
        vals[k] += 0.75
        reg.set_profile(1, 1, side, var, vals)
        out = fill(reg)
        for a, b in zip(out, base):
            jj, ii = np.nonzero(a != b)
            line = jj if side in (case.WEST, case.EAST) else ii
            # a later side's fill may read a changed value at a corner
            corner = np.isin(ii, (0, 1, nx, nx + 1)) & np.isin(jj, (0, 1, ny, ny + 1))
            assert np.all((line == k) | corner), (k, jj, ii)
            if jj.size:
                moved.add(k)
    # the (side, variable) pairs whose boundary type reads the value do react
    used = {(case.WEST, case.VAR_U), (case.WEST, case.VAR_V), (case.WEST, case.VAR_P),
            (case.SOUTH, case.VAR_U), (case.NORTH, case.VAR_U), (case.SOUTH, case.VAR_P),
            (case.NORTH, case.VAR_P), (case.EAST, case.VAR_U), (case.EAST, case.VAR_P)}
    if (side, var) in used:
        assert moved, "the profile is never read"


def test_side_positions_on_a_uniform_grid():
    nx, ny = 9, 7
    x, y = case.uniform_nodes(nx, ny, lx=2.0, ly=3.0)
    reg = case.cavity_regions(nx, ny)
    j = np.arange(ny + 2)
    i = np.arange(nx + 2)
    for side in (case.WEST, case.EAST):
        # u and p at cell centres (between nodes j-1 and j), v on node j
        assert np.allclose(case.side_positions(x, y, reg, 1, 1, side, case.VAR_U), (j - 1.5) / (ny - 1))
        assert np.allclose(case.side_positions(x, y, reg, 1, 1, side, case.VAR_P), (j - 1.5) / (ny - 1))
        assert np.allclose(case.side_positions(x, y, reg, 1, 1, side, case.VAR_V), (j - 1.0) / (ny - 1))
    for side in (case.SOUTH, case.NORTH):
        assert np.allclose(case.side_positions(x, y, reg, 1, 1, side, case.VAR_V), (i - 1.5) / (nx - 1))
        assert np.allclose(case.side_positions(x, y, reg, 1, 1, side, case.VAR_U), (i - 1.0) / (nx - 1))


def test_side_positions_of_an_inner_region():
    nx, ny = 11, 9
    x, y = case.uniform_nodes(nx, ny)
    reg = case.make_regions(nx, ny, nreg=(2, 2), i_borders=(5,), j_borders=(4,))
    s = case.side_positions(x, y, reg, 2, 2, case.WEST, case.VAR_V)
    assert reg.border(2, 2, case.SOUTH) == 4 and reg.border(2, 2, case.NORTH) == ny
    assert s[4] == 0.0 and s[ny] == 1.0


def test_parabolic_inflow_flux_closed_form():
    """The midpoint rule over the cell-centred u values of 4 s (1 - s) gives
    2/3 + h^2/3 exactly (h the cell size in s)."""
    nx, ny = 17, 33
    x, y = case.uniform_nodes(nx, ny, lx=4.0)
    reg = case.cavity_regions(nx, ny)
    s = case.side_positions(x, y, reg, 1, 1, case.WEST, case.VAR_U)
    u = case.parabolic_profile(s, 1.0)
    h = 1.0 / (ny - 1)
    assert abs(h * u[2:ny + 1].sum() - (2.0 / 3.0 + h * h / 3.0)) < 1e-14
    # the peak sits at the middle node line, and the wall nodes carry 0
    sv = case.side_positions(x, y, reg, 1, 1, case.WEST, case.VAR_V)
    v = case.parabolic_profile(sv, 1.0)
    assert v[1] == 0.0 and v[ny] == 0.0 and v[(ny + 1) // 2] == 1.0


def test_sine_profile_modes():
    s = np.linspace(0, 1, 101)
    a = [0.3, -0.1, 0.05]
    f = case.sine_profile(s, a)
    assert abs(f[0]) < 1e-15 and abs(f[-1]) < 1e-15
    # orthogonality: 2 * integral f sin(k pi s) ds = a_k (trapezoid, exact for these modes)
    for k, ak in enumerate(a, start=1):
        w = np.sin(k * np.pi * s) * f
        assert abs(2 * np.trapezoid(w, s) - ak) < 1e-12


def test_set_profile_validates():
    reg = case.cavity_regions(9, 7)
    with pytest.raises(ValueError):
        reg.set_profile(1, 1, case.WEST, case.VAR_U, np.zeros(9 + 2))   # west needs ny+2
    with pytest.raises(ValueError):
        reg.set_profile(2, 1, case.WEST, case.VAR_U, np.zeros(7 + 2))
    with pytest.raises(ValueError):
        reg.set_profile(1, 1, case.WEST, 4, np.zeros(7 + 2))
    reg.set_profile(1, 1, case.NORTH, case.VAR_U, np.ones(11))
    poff, pval = reg.profile_tables()
    assert poff.shape == (reg.mgri, reg.mgrj, 4, 3) and poff.flags.f_contiguous
    assert poff[0, 0, case.NORTH - 1, case.VAR_U - 1] == 0 and (poff >= 0).sum() == 1 and pval.size == 11
    reg.set_profile(1, 1, case.NORTH, case.VAR_U, None)
    assert reg.profile_tables() == (None, None)


def _channel_deck(tmp_path, extra):
    fx = helpers.load("run_chan_64x16.npz")
    import refdecks
    refdecks.write_grid(str(tmp_path / "grid.xyz"), fx["x"], fx["y"])
    refdecks.write_fluid_table(str(tmp_path / "fluid.dat"), "fl01", float(fx["rho"]), float(fx["mu"]))
    text = str(fx["deck"]).replace("outlet 1 1 east mass_cons", "outlet 1 1 east mass_cons\n" + extra)
    (tmp_path / "input.dat").write_text(text)
    d = deck.read_input(str(tmp_path / "input.dat"))
    x, y = deck.read_grid(str(tmp_path / "grid.xyz"))
    return fx, d, x, y


def test_deck_bc_profile_forms(tmp_path):
    ny = 17
    vals = np.linspace(0.0, 1.0, ny + 2)
    (tmp_path / "lid.dat").write_text("\n".join(repr(float(a)) for a in vals))
    fx, d, x, y = _channel_deck(tmp_path, "bc_profile 1 1 west u parabolic 1.5\n"
                                          "bc_profile 1 1 east p file lid.dat\n"
                                          "bc_profile 1 1 north u sine 0.1 -0.2")
    reg = deck.build_case(d, float(fx["rho"]), 65, ny, x=x, y=y, case_dir=str(tmp_path))
    s = case.side_positions(x, y, reg, 1, 1, case.WEST, case.VAR_U)
    assert np.array_equal(reg.profiles[(1, 1, case.WEST, case.VAR_U)], case.parabolic_profile(s, 1.5))
    rho = float(fx["rho"])
    assert np.array_equal(reg.profiles[(1, 1, case.EAST, case.VAR_P)], vals / (rho * 1.0 * 1.0))
    s = case.side_positions(x, y, reg, 1, 1, case.NORTH, case.VAR_U)
    assert np.array_equal(reg.profiles[(1, 1, case.NORTH, case.VAR_U)], case.sine_profile(s, [0.1, -0.2]))
    with pytest.raises(deck.DeckError, match="needs the grid"):
        deck.build_case(d, rho, 65, ny)


def test_deck_bc_profile_errors(tmp_path):
    (tmp_path / "short.dat").write_text("1 2 3")
    fx, d, x, y = _channel_deck(tmp_path, "bc_profile 1 1 west u file short.dat")
    with pytest.raises(deck.DeckError, match="needs 19 values"):
        deck.build_case(d, float(fx["rho"]), 65, 17, x=x, y=y, case_dir=str(tmp_path))
    with pytest.raises(deck.DeckError):
        _channel_deck(tmp_path, "bc_profile 1 1 west w parabolic 1.0")
