"""Thermal routines on the GPU (wd2_temp_bound_cond, wd2_therm_energy,
wd2_eq_state, wd2_filter) against the reference's goldens and the oracle,
bit for bit."""
import numpy as np
import pytest

import helpers
from test_thermal import THERMAL_FIXTURES, thermal_from
from wolfd2_amd import api, case
from wolfd2_amd._lib import Wd2Error

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", THERMAL_FIXTURES)
def test_thermal_bitwise_vs_reference_goldens(name):
    fx = helpers.load(name)
    nx, ny = int(fx["nx"]), int(fx["ny"])
    print(f"thermal goldens {name}", flush=True)
    g = case.compute_grid(fx["x"], fx["y"])
    reg = helpers.regions_from(fx)
    th = thermal_from(fx)
    E = lambda a: helpers.embed(a, g.ld, g.ncol)  # noqa: E731
    S = np.s_[: ny + 2, : nx + 2]
    t = E(fx["tbc_in"])
    api.temp_bound_cond(reg, th, t)
    assert np.array_equal(t[S], fx["tbc_out"])
    t = E(fx["te_t"])
    api.therm_energy(reg, th, g, float(fx["te_dk"]), float(fx["te_pe"]), E(fx["te_un"]), E(fx["te_vn"]),
                     E(fx["te_u"]), E(fx["te_v"]), E(fx["te_tn"]), t)
    assert np.array_equal(t[S], fx["te_out"])
    den = np.zeros_like(t)
    api.eq_state(g, *fx["eqs_args"], E(fx["eqs_p"]), E(fx["eqs_t"]), den)
    assert np.array_equal(den[S], fx["eqs_out"])
    for comp, nm in ((1, "u"), (2, "v"), (4, "t")):
        q = E(fx[f"filt_{nm}_in"])
        api.filter(reg, th, comp, 20.0, q)
        assert np.array_equal(q[S], fx[f"filt_{nm}_out"]), nm


@pytest.mark.parametrize("nx,ny", [(130, 97), (1025, 1025)])
def test_thermal_bitwise_vs_oracle(oracle, nx, ny):
    """A multi-region layout with every thermal region and boundary type; at
    1025^2 nodes the energy chains are 1,048,576 rows long."""
    print(f"thermal vs oracle {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx)
    x, y = case.skewed_nodes(nx, ny, seed=4) if nx < 500 else case.uniform_nodes(nx, ny)
    g = case.compute_grid(x, y)
    ib, jb = (nx // 3, 2 * nx // 3), (ny // 2,)
    reg = case.make_regions(nx, ny, nreg=(3, 2), i_borders=ib, j_borders=jb,
                            walls=[(2, 2, "north", "no_slip", 1.0, 0.0)],
                            outlets=[(3, 1, "east", "fully_dev", 0.0, 0.0, 0.0)])
    th = case.make_thermal(reg, heat=[(2, 1, 2.0)], fixed=[(1, 2, 0.4)],
                           walls=[(1, 1, "west", "temper", 1.0), (3, 2, "north", "flux", 0.01)])

    def rf(s=1.0):
        return s * rng.standard_normal((g.ncol, g.ld))
    un, vn, u, v, tn = rf(0.3), rf(0.3), rf(0.3), rf(0.3), rf()
    tg = rf()
    to = tg.copy()
    for _ in range(2):      # two calls: the second starts from the first's t
        api.therm_energy(reg, th, g, 1e-3, 200.0, un, vn, u, v, tn, tg)
        oracle.therm_energy(reg, th, g, 1e-3, 200.0, un, vn, u, v, tn, to)
        assert np.array_equal(tg, to)
    for comp in (1, 2, 4):
        qg = rf()
        qo = qg.copy()
        api.filter(reg, th, comp, 12.0, qg)
        oracle.filter(reg, th, g.ld, comp, 12.0, qo)
        assert np.array_equal(qg, qo), comp
    p, t = rf(0.2), rf(0.5)
    dg, do = rf(), None
    do = dg.copy()
    api.eq_state(g, 1.5, 1.2, 400.0, 300.0, 287.0, p, t, dg)
    oracle.eq_state(g, 1.5, 1.2, 400.0, 300.0, 287.0, p, t, do)
    assert np.array_equal(dg, do)
    tg, to = rf(), None
    to = tg.copy()
    api.temp_bound_cond(reg, th, tg)
    oracle.temp_bound_cond(reg, th, g.ld, to)
    assert np.array_equal(tg, to)


def test_thermal_bad_flags_raise():
    nx = ny = 17
    x, y = case.uniform_nodes(nx, ny)
    g = case.compute_grid(x, y)
    reg = case.make_regions(nx, ny)
    th = case.make_thermal(reg)
    th.nTemBdTp[0, 0, 0] = 7
    t = np.zeros((g.ncol, g.ld))
    with pytest.raises(Wd2Error, match="nTemBdTp"):
        api.temp_bound_cond(reg, th, t)
    with pytest.raises(Wd2Error, match="ncomp"):
        api.filter(reg, case.make_thermal(reg), 3, 4.0, t)
    # the context recovers: a good call after the errors works
    th = case.make_thermal(reg, walls=[(1, 1, "west", "temper", 1.0)])
    api.temp_bound_cond(reg, th, t)
    assert np.all(t[2:ny + 1, 1] == 2.0)
