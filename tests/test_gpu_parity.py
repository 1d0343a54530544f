"""Parity of the HIP path with the reference (through the C ABI).

FP64 results must be bit-identical to the reference's: every comparison here
is np.array_equal (tolerance 0) against the golden fixtures the reference
wrote, or against the oracle (the C restatement) on the same seeded inputs.
"""
import numpy as np
import pytest

import helpers
from wolfd2_amd import api, case

pytestmark = pytest.mark.gpu


def _kernel_case(fx):
    g = case.compute_grid(fx["x"], fx["y"])
    return int(fx["nx"]), int(fx["ny"]), g, helpers.regions_from(fx)


@pytest.mark.parametrize("name", helpers.kernel_fixture_names())
def test_kernels_bitwise_vs_reference_goldens(name):
    fx = helpers.load(name)
    nx, ny, g, reg = _kernel_case(fx)
    ld, nc = g.ld, g.ncol
    E = lambda a: helpers.embed(a, ld, nc)  # noqa: E731
    S = np.s_[: ny + 2, : nx + 2]
    print(f"case {name} nx={nx} ny={ny}", flush=True)

    u, v = E(fx["vbc_u_in"]), E(fx["vbc_v_in"])
    api.vel_bound_cond(reg, u, v)
    assert np.array_equal(u[S], fx["vbc_u_out"]) and np.array_equal(v[S], fx["vbc_v_out"])
    u, v = E(fx["vob_u_in"]), E(fx["vob_v_in"])
    api.vel_outflow_bcs(reg, u, v)
    assert np.array_equal(u[S], fx["vob_u_out"]) and np.array_equal(v[S], fx["vob_v_out"])
    p = E(fx["pbc_p_in"])
    api.pres_bound_cond(reg, p)
    assert np.array_equal(p[S], fx["pbc_p_out"])
    u, v, p = E(fx["prj_u_in"]), E(fx["prj_v_in"]), E(fx["prj_p"])
    api.project(reg, g, float(fx["prj_dk"]), p, u, v)
    assert np.array_equal(u[S], fx["prj_u_out"]) and np.array_equal(v[S], fx["prj_v_out"])
    div = np.zeros_like(u)
    api.divergence(g, 1, E(fx["div_u"]), E(fx["div_v"]), div)
    assert np.array_equal(div[S], fx["div_out"])
    us, vs, un, vn = (E(fx[f"mom_{k}"]) for k in ("us", "vs", "un", "vn"))
    prm = case.Params(dk=0.01, re=100.0)
    dus, dvs, d0 = np.zeros_like(us), np.zeros_like(us), np.zeros_like(us)
    api.x_momentum(reg, g, prm, us, vs, un, vn, dus)
    api.y_momentum(reg, g, prm, d0, d0.copy(), us, vs, un, vn, dvs)
    assert np.array_equal(dus[S], fx["mom_dus"]), "XMomentum"
    assert np.array_equal(dvs[S], fx["mom_dvs"]), "YMomentum"
    for mq, qt in ((1, 1e-4), (30, 1e-10)):
        us2, vs2 = np.zeros_like(us), np.zeros_like(us)
        n = api.naux_momentum(reg, g, prm, d0, d0.copy(), un, vn, us2, vs2, mqiter=mq, qtol=qt)
        assert n == int(fx[f"naux{mq}_n"])
        assert np.array_equal(us2[S], fx[f"naux{mq}_us"]) and np.array_equal(vs2[S], fx[f"naux{mq}_vs"])
    assert api.dmax_norm(nx, ny, dus) == float(fx["dmax"])
    assert api.diff_max_norm(nx, ny, us, vs) == float(fx["diffmax"])
    x = api.alt_trid_lu(fx["trid_a"], fx["trid_b"])
    assert np.array_equal(x, fx["trid_x"])


@pytest.mark.parametrize("name", helpers.kernel_fixture_names())
@pytest.mark.parametrize("solver", [1, 2, 3, 4, 5, 6])
def test_ppe_solvers_bitwise_vs_reference_goldens(name, solver):
    fx = helpers.load(name)
    nx, ny, g, reg = _kernel_case(fx)
    E = lambda a: helpers.embed(a, g.ld, g.ncol)  # noqa: E731
    S = np.s_[: ny + 2, : nx + 2]
    for cart in (1, 0):
        key = f"ppe{solver}c{cart}"
        print(f"case {name} {key}", flush=True)
        p = E(fx[f"{key}_p_in"])
        prm = case.Params(dk=0.01, re=100.0, nPpeSolver=solver, msorit=25, sortol=1e-8, sorrel=1.3,
                          cartesian=bool(cart))
        n = api.ppe(reg, g, prm, E(fx[f"{key}_u"]), E(fx[f"{key}_v"]), p)
        assert n == int(fx[f"{key}_n"])
        assert np.array_equal(p[S], fx[f"{key}_p_out"]), key


@pytest.mark.parametrize("name", helpers.kernel_fixture_names())
def test_lex_sor_tiled_wavefront_path_bitwise(name):
    """Sor on the tiled wavefront (one launch per sweep) instead of the
    single-workgroup row-per-lane kernel: the same reference goldens."""
    from wolfd2_amd._lib import lib
    fx = helpers.load(name)
    nx, ny, g, reg = _kernel_case(fx)
    E = lambda a: helpers.embed(a, g.ld, g.ncol)  # noqa: E731
    S = np.s_[: ny + 2, : nx + 2]
    old = lib().wd2_set_lex_rows(0)
    try:
        for cart in (1, 0):
            key = f"ppe1c{cart}"
            p = E(fx[f"{key}_p_in"])
            prm = case.Params(dk=0.01, re=100.0, nPpeSolver=1, msorit=25, sortol=1e-8, sorrel=1.3,
                              cartesian=bool(cart))
            n = api.ppe(reg, g, prm, E(fx[f"{key}_u"]), E(fx[f"{key}_v"]), p)
            assert n == int(fx[f"{key}_n"])
            assert np.array_equal(p[S], fx[f"{key}_p_out"]), key
    finally:
        lib().wd2_set_lex_rows(old)


def _chain(n, rng, weak=False):
    a = np.zeros((n, 3))
    a[:, 0] = rng.uniform(-1, 0, n)
    a[:, 2] = rng.uniform(-1, 0, n)
    a[:, 1] = (np.abs(a[:, 0]) + np.abs(a[:, 2])) * (1.0001 if weak else 1.3) + (0 if weak else 0.05)
    b = rng.standard_normal(n)
    return a, b


@pytest.mark.parametrize("n,weak", [(2, False), (3, False), (17, False), (4095, False), (4096, False),
                                    (4097, False), (100_003, False), (1_049_600, False), (20_000, True)])
def test_alt_trid_lu_bitwise_vs_oracle(oracle, n, weak):
    rng = np.random.default_rng(n)
    a, b = _chain(n, rng, weak)
    print(f"chain n={n} weak={weak}", flush=True)
    x_gpu = api.alt_trid_lu(a, b)
    _, x_ref = oracle.alt_trid_lu(a, b)
    assert np.array_equal(x_gpu, x_ref)


def _two_fixed_points(rng):
    """Constant rows (a, d, c, b) whose rounded forward recurrence has two
    fixed points (a run from the zero state and a run from another state
    settle on different values and never coincide bit for bit), and whose
    backward recurrence x = b' - c' x on either settled (c', b') has two
    fixed points as well."""
    def run(a, d, c, b, cp, bp, n):
        out = []
        for _ in range(n):
            dp = d - (a * cp)
            cp = c / dp
            bp = (b - a * bp) / dp
            out.append((cp, bp))
        return out

    def back(cp, bp, x, n):
        out = []
        for _ in range(n):
            x = bp - (cp * x)
            out.append(x)
        return out

    def apart(r1, r2):
        return all(p != q for p, q in zip(r1, r2))
    while True:
        a, c = -rng.uniform(0.05, 0.5), -rng.uniform(0.05, 0.5)
        d, b = 1.0 + rng.uniform(0.0, 0.5), rng.uniform(-1.0, 1.0)
        r1, r2 = run(a, d, c, b, 0.0, 0.0, 200), run(a, d, c, b, -0.3, 0.7, 200)
        if apart(r1, r2) and all(apart(back(*r[-1], 0.0, 100), back(*r[-1], 3.0, 100)) for r in (r1, r2)):
            return a, d, c, b


def test_alt_trid_lu_rounding_cycles_bitwise(oracle):
    """Stretches of identical rows on which speculative runs never coincide
    with the exact one (two fixed points of the rounded recurrence): inside a
    2048-row block, across block boundaries and longer than a block.  The
    fix-up rounds' pass through constant-even segments (forward and backward)
    and the cross-block repairs must still give the sequential sweep's bits."""
    rng = np.random.default_rng(7)
    a, b = _chain(30_000, rng)
    row = _two_fixed_points(rng)
    for k0, length in [(100, 300), (1500, 1200), (5000, 40), (7000, 2600), (12_000, 5000), (20_000, 900)]:
        a[k0:k0 + length] = row[:3]
        b[k0:k0 + length] = row[3]
    print("chain with rounding-cycle stretches", flush=True)
    x_gpu = api.alt_trid_lu(a, b)
    _, x_ref = oracle.alt_trid_lu(a, b)
    assert np.array_equal(x_gpu, x_ref)


def test_alt_trid_lu_sequential_fallback_bitwise(oracle):
    """The 1-D Laplacian chain (d = 2, a = c = -1) barely contracts, so the
    speculative segments never coincide bit for bit with the repaired ones:
    the solver must detect it and fall back to the sequential sweep."""
    from wolfd2_amd._lib import lib
    n = 5000
    a = np.zeros((n, 3))
    a[:, 0], a[:, 1], a[:, 2] = -1.0, 2.0, -1.0
    b = np.random.default_rng(11).standard_normal(n)
    before = lib().wd2_trid_fallbacks()
    x_gpu = api.alt_trid_lu(a, b)
    after = lib().wd2_trid_fallbacks()
    _, x_ref = oracle.alt_trid_lu(a, b)
    assert np.array_equal(x_gpu, x_ref)
    assert after > before, "expected the sequential fallback to run"


@pytest.fixture(params=[1, 0], ids=["tiled_sor", "colour_sor"])
def sor_tiling(request):
    from wolfd2_amd._lib import lib
    old = lib().wd2_set_tiled_sor(request.param)
    yield request.param
    lib().wd2_set_tiled_sor(old)


@pytest.mark.parametrize("name", helpers.run_fixture_names())
def test_whole_runs_bitwise_vs_reference_program(name, sor_tiling):
    fx = helpers.load(name)
    if not sor_tiling and int(fx["ppe_solver"]) not in (5, 6):
        pytest.skip("the tiling switch only changes the red-black SOR path")
    g, reg, prm = helpers.run_case(fx)
    nx, ny = g.nx, g.ny
    print(f"run {name} nx={nx} ny={ny} solver={prm.nPpeSolver} steps={int(fx['nts'])}", flush=True)
    sol = api.Solver(g, reg, prm)
    z = np.zeros((g.ncol, g.ld))
    sol.set_fields(z, z, z)
    sol.cold_start()
    table = sol.step(int(fx["nts"]))
    u, v, p = sol.get_fields()
    sol.close()
    ref = fx["table"]
    n = min(len(ref), len(table))
    assert np.array_equal(table[:n, 0], ref[:n, 2]), "QlIt differs"
    assert np.array_equal(table[:n, 1], ref[:n, 3]), "PpeIt differs"
    fin = np.isfinite(ref[:n, 4:7])
    np.testing.assert_allclose(table[:n, 2:5][fin], ref[:n, 4:7][fin], rtol=6e-6, atol=1e-300)
    S = np.s_[: ny + 2, : nx + 2]
    if np.all(np.isfinite(fx["u"])):
        assert np.array_equal(u[S], fx["u"]), "u"
        assert np.array_equal(v[S], fx["v"]), "v"
        assert np.array_equal(p[S], fx["p"]), "p"


@pytest.mark.parametrize("solver", [5, 1])
def test_full_size_cavity_steps_bitwise_vs_oracle(oracle, solver, sor_tiling):
    """1024^2 cells (the benchmark grid): a cold start and two time steps with
    a capped Poisson solve, GPU against the oracle, bit for bit."""
    nx = ny = 1025
    x, y = case.uniform_nodes(nx, ny)
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(nx, ny)
    if not sor_tiling and solver != 5:
        pytest.skip("the tiling switch only changes the red-black SOR path")
    # 13 sweeps: three tiled launches of four and a one-sweep tail
    prm = case.Params.from_physical(2e-4, 1.0, 1e-3, nPpeSolver=solver, msorit=13 if solver == 5 else 4)
    print(f"full-size cavity solver={solver} tiled={sor_tiling}", flush=True)
    sol = api.Solver(g, reg, prm)
    z = np.zeros((g.ncol, g.ld))
    sol.set_fields(z, z, z)
    sol.cold_start()
    tg = sol.step(2)
    u1, v1, p1 = sol.get_fields()
    sol.close()
    u, v, p = (np.zeros((g.ncol, g.ld)) for _ in range(3))
    oracle.cold_start(reg, g, prm, u, v, p)
    to = oracle.run(reg, g, prm, u, v, p, 2)
    assert np.array_equal(tg[:, :2], to[:, :2])
    assert np.array_equal(u1, u) and np.array_equal(v1, v) and np.array_equal(p1, p)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_uniform_coefficient_path_equals_array_path(precision):
    """On a uniform Cartesian grid the tiled red-black SOR kernels take rau and
    rgv as two values; loading the arrays instead gives the same bits."""
    from wolfd2_amd._lib import lib
    nx = ny = 1025
    x, y = case.uniform_nodes(nx, ny)
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(nx, ny)
    prm = case.Params.from_physical(2e-4, 1.0, 1e-3, nPpeSolver=5, msorit=60)
    print(f"uniform vs array coefficients {precision}", flush=True)
    out = []
    old = lib().wd2_set_uniform_coefficients(1)
    try:
        for flag in (1, 0):
            lib().wd2_set_uniform_coefficients(flag)
            sol = api.Solver(g, reg, prm)
            sol.set_precision(precision)
            z = np.zeros((g.ncol, g.ld))
            sol.set_fields(z, z, z)
            sol.cold_start()
            tab = sol.step(2)
            out.append((tab, *sol.get_fields()))
            sol.close()
    finally:
        lib().wd2_set_uniform_coefficients(old)
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_stretched_cartesian_grid_tiled_sor_vs_oracle(oracle):
    """A stretched Cartesian grid (rau, rgv vary, so the kernels load them):
    tiled red-black SOR steps against the oracle, bit for bit."""
    nx, ny = 161, 129
    s = np.arange(nx) / (nx - 1.0)
    t = np.arange(ny) / (ny - 1.0)
    xs = s + 0.12 * np.sin(np.pi * s) * s
    yt = t - 0.08 * np.sin(2 * np.pi * t)
    x = np.broadcast_to(xs[None, :], (ny, nx)).copy()
    y = np.broadcast_to(yt[:, None], (ny, nx)).copy()
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(nx, ny)
    prm = case.Params.from_physical(5e-4, 1.0, 1e-2, nPpeSolver=5, msorit=200, sorrel=1.7, cartesian=True)
    print("stretched Cartesian grid, tiled SOR", flush=True)
    sol = api.Solver(g, reg, prm)
    z = np.zeros((g.ncol, g.ld))
    sol.set_fields(z, z, z)
    sol.cold_start()
    tg = sol.step(3)
    fg = sol.get_fields()
    sol.close()
    u, v, p = (np.zeros((g.ncol, g.ld)) for _ in range(3))
    oracle.cold_start(reg, g, prm, u, v, p)
    to = oracle.run(reg, g, prm, u, v, p, 3)
    assert np.array_equal(tg, to)
    for a, b in zip(fg, (u, v, p)):
        assert np.array_equal(a, b)


def _stretched_case(nx=161, ny=129, **kw):
    s = np.arange(nx) / (nx - 1.0)
    t = np.arange(ny) / (ny - 1.0)
    xs = s + 0.12 * np.sin(np.pi * s) * s
    yt = t - 0.08 * np.sin(2 * np.pi * t)
    x = np.broadcast_to(xs[None, :], (ny, nx)).copy()
    y = np.broadcast_to(yt[:, None], (ny, nx)).copy()
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(nx, ny)
    prm = case.Params.from_physical(5e-4, 1.0, 1e-2, cartesian=True, **kw)
    return g, reg, prm


def test_stretched_cartesian_grid_mixed_follows_fp64():
    """The mixed kernel's per-point coefficient path (rau, rgv loaded per
    point, distinct a1/a5 and a2/a4 at every point of a stretched grid): a
    swapped or shifted coefficient changes the fixed point, which the FP64
    path (bit-exact against the oracle above) pins.  Bar: sweep counts within
    one of FP64's, fields within 1e-6 relative L2."""
    g, reg, prm = _stretched_case(nPpeSolver=5, msorit=200, sorrel=1.7)
    print("stretched Cartesian grid, mixed vs fp64", flush=True)
    res = {}
    for prec in ("fp64", "mixed"):
        sol = api.Solver(g, reg, prm)
        sol.set_precision(prec)
        z = np.zeros((g.ncol, g.ld))
        sol.set_fields(z, z, z)
        sol.cold_start()
        res[prec] = (sol.step(3), sol.get_fields())
        sol.close()
    (t64, f64), (tmx, fmx) = res["fp64"], res["mixed"]
    assert np.array_equal(t64[:, 0], tmx[:, 0])
    assert np.all(np.abs(t64[:, 1] - tmx[:, 1]) <= 1)
    S = np.s_[: g.ny + 2, : g.nx + 2]
    for name, a, b in zip("uvp", fmx, f64):
        e = np.linalg.norm(a[S] - b[S]) / np.linalg.norm(b[S])
        print(f"  {name}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-6, name


@pytest.mark.parametrize("grid", ["uniform", "skewed"])
def test_lex_rows_several_wavefronts_bitwise_vs_oracle(oracle, grid):
    """Lexicographic Sor as one single-workgroup launch with 300 pressure rows
    (five wavefronts: the cross-wavefront ring hand-off, the skew between
    wavefronts and the barriers between them), on a uniform Cartesian grid
    and on a skewed non-Cartesian one (RhsPpe inside the sweeps)."""
    nx, ny = 161, 301
    if grid == "uniform":
        x, y = case.uniform_nodes(nx, ny, lx=0.6)
        cart = True
    else:
        x, y = case.skewed_nodes(nx, ny, seed=5)
        x = 0.55 * x   # cells about as wide as they are high
        cart = False
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(nx, ny)
    prm = case.Params.from_physical(2e-4, 1.0, 1e-2, nPpeSolver=1, msorit=1500, sorrel=1.8, cartesian=cart)
    print(f"lexicographic Sor, {ny - 1} rows, {grid}", flush=True)
    from wolfd2_amd._lib import lib
    old = lib().wd2_set_lex_rows(1)
    try:
        sol = api.Solver(g, reg, prm)
        z = np.zeros((g.ncol, g.ld))
        sol.set_fields(z, z, z)
        sol.cold_start()
        tg = sol.step(2)
        fg = sol.get_fields()
        sol.close()
    finally:
        lib().wd2_set_lex_rows(old)
    u, v, p = (np.zeros((g.ncol, g.ld)) for _ in range(3))
    oracle.cold_start(reg, g, prm, u, v, p)
    to = oracle.run(reg, g, prm, u, v, p, 2)
    assert np.all(np.isfinite(to)) and np.all(np.abs(u) < 10.0), "the case must stay bounded to compare"
    assert np.array_equal(tg, to)
    for a, b in zip(fg, (u, v, p)):
        assert np.array_equal(a, b)
