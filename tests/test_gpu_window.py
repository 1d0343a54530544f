"""The mixed path's windowed FP32 chain solver (partitioned Thomas + PCR in
LDS over overlapping windows, wd2_set_window_chains) against the bit-exact
FP32 speculative solver and the FP64 path.  In a time step the window solver
serves the q-l iterations after the first; the first keeps the sequential
solver.

* Where every window is diagonally dominant (dominance ratio <= 0.9, the
  cavities and channels of the measured workloads), the two FP32 solvers agree
  to FP32 rounding: after tens of steps the two mixed runs differ by less than
  1e-6 relative L2 (u and v against the norm of the velocity, p against its
  own), and the window solver's run is no further from FP64 than the exact
  solver's (within 10 %); on the cavity both are within 1e-6 of FP64.
* On random batches of chains the window solver agrees with an FP64 Thomas
  solve of the same matrices (first-row quirk included) to FP32 accuracy.
  Where a window is not dominant enough (ratio 0.95 against the solver's
  0.9) the bit-exact FP32 solver takes over and x is its result bit for bit.
"""
import numpy as np
import pytest

from wolfd2_amd import api, case, workloads
from wolfd2_amd._lib import lib

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _run(g, reg, prm, precision, window, nsteps, fields=None):
    old = lib().wd2_set_window_chains(1 if window else 0)
    try:
        sol = api.Solver(g, reg, prm)
        sol.set_precision(precision)
        if fields is None:
            z = np.zeros((g.ncol, g.ld))
            sol.set_fields(z, z, z)
            sol.cold_start()
        else:
            sol.set_fields(*fields)
        tab = sol.step(nsteps)
        out = sol.get_fields()
        sol.close()
    finally:
        lib().wd2_set_window_chains(old)
    return out, tab


@pytest.mark.parametrize("case_name", ["cavity_256", "channel_512x128", "stretched_cavity"])
def test_window_chains_agree_with_the_exact_fp32_solver(case_name):
    if case_name == "cavity_256":
        g, reg, prm = workloads.cavity(256, 1000.0, 2e-4, nPpeSolver=7, msorit=100, sortol=1e-10)
        nsteps, fields = 40, None
    elif case_name == "channel_512x128":
        g, reg, prm, x, y = workloads.channel(512, 128, outlet="mass_cons", nPpeSolver=7, msorit=100, sortol=1e-10)
        nsteps, fields = 30, workloads.poiseuille_noise(g, reg, x, y)
    else:
        nx, ny = 161, 129
        s = np.arange(nx) / (nx - 1.0)
        t = np.arange(ny) / (ny - 1.0)
        xs = s + 0.12 * np.sin(np.pi * s) * s
        yt = t - 0.08 * np.sin(2 * np.pi * t)
        x = np.broadcast_to(xs[None, :], (ny, nx)).copy()
        y = np.broadcast_to(yt[:, None], (ny, nx)).copy()
        g = case.compute_grid(x, y)
        reg = case.cavity_regions(g.nx, g.ny)
        prm = case.Params.from_physical(5e-4, 1.0, 1.0 / 400.0, nPpeSolver=7, msorit=100, sortol=1e-10)
        nsteps, fields = 30, None
    print(f"  {case_name}: mixed with the window solver, the exact FP32 solver, FP64", flush=True)
    fw, tw = _run(g, reg, prm, "mixed", True, nsteps, fields)
    fe, te = _run(g, reg, prm, "mixed", False, nsteps, fields)
    f64, _ = _run(g, reg, prm, "fp64", True, nsteps, fields)
    S = np.s_[: g.ny + 2, : g.nx + 2]
    # velocity errors against the norm of the velocity (u, v) (the channel's
    # v is the 1e-3 noise alone), p's against its own norm (as the channel
    # test of the measured workloads)
    speed = np.sqrt(np.linalg.norm(f64[0][S]) ** 2 + np.linalg.norm(f64[1][S]) ** 2)
    for name, a, b, c in zip("uvp", fw, fe, f64):
        if name == "p":
            a, b, c = a[S], b[S], c[S]
            d_we, d_w64, d_e64 = rel_l2(a, b), rel_l2(a, c), rel_l2(b, c)
        else:
            a, b, c = a[S], b[S], c[S]
            d_we, d_w64, d_e64 = (np.linalg.norm(x - y) / speed for x, y in ((a, b), (a, c), (b, c)))
        print(f"    {name}: window vs exact {d_we:.2e}; vs FP64: window {d_w64:.2e}, exact {d_e64:.2e}", flush=True)
        # the window solver adds nothing that matters to the mixed path's error
        assert d_we < 1e-6
        assert d_w64 <= 1.1 * d_e64 + 1e-8
        if case_name == "cavity_256":   # the measured configuration's bar
            assert d_w64 < 1e-6
    print(f"    q-l iterations window {tw[:, 0].mean():.2f}, exact {te[:, 0].mean():.2f}", flush=True)


def _chains(rng, nchain, L, ratio):
    """nchain chains of L rows with |a| + |c| = ratio |d| (random signs and
    splits), float32; the first row's a and the last row's c are junk the
    solver must ignore."""
    n = nchain * L
    d = rng.uniform(1.0, 3.0, n) * rng.choice([-1.0, 1.0], n)
    split = rng.uniform(0.0, 1.0, n)
    a = ratio * np.abs(d) * split * rng.choice([-1.0, 1.0], n)
    c = ratio * np.abs(d) * (1.0 - split) * rng.choice([-1.0, 1.0], n)
    b = rng.standard_normal(n)
    return [q.astype(np.float32) for q in (a, d, c, b)]


def _thomas64(a, d, c, b, L):
    """FP64 AltTridLU of every chain of L rows (first-row quirk: row 1's
    super-diagonal is divided by row 2's diagonal)."""
    A, D, Cc, B = (q.astype(np.float64).reshape(-1, L) for q in (a, d, c, b))
    cp = np.zeros_like(D)
    bp = np.zeros_like(D)
    if L > 1:
        cp[:, 0] = Cc[:, 0] / D[:, 1]
    bp[:, 0] = B[:, 0] / D[:, 0]
    for k in range(1, L):
        den = D[:, k] - A[:, k] * cp[:, k - 1]
        if k < L - 1:
            cp[:, k] = Cc[:, k] / den
        bp[:, k] = (B[:, k] - A[:, k] * bp[:, k - 1]) / den
    x = bp.copy()
    for k in range(L - 2, -1, -1):
        x[:, k] = bp[:, k] - cp[:, k] * x[:, k + 1]
    return x.reshape(-1)


def _solve32(a, d, c, b, n1, L1, L2, window):
    import ctypes as C
    f = lambda q: q.ctypes.data_as(C.POINTER(C.c_float))
    x = np.zeros_like(b)
    rc = lib().wd2_trid_f32(C.c_long(b.size), C.c_long(n1), C.c_long(L1), C.c_long(L2), f(a), f(d), f(c), f(b), f(x),
                            int(window))
    assert rc == 0
    return x


@pytest.mark.parametrize("L1,L2,nch1,nch2", [(263169, 262656, 1, 1), (1000, 777, 300, 400), (4, 3, 1000, 1000)])
def test_window_solver_on_random_dominant_chains(L1, L2, nch1, nch2):
    """Batches of two chain lengths (one long chain each as the momentum
    batches have, or many short ones), dominance ratio 0.6: the window
    solver's x agrees with an FP64 Thomas solve of the same quirked matrices
    to FP32 accuracy, as the bit-exact FP32 solver does, in each window shape
    (128, 256 and 512 lanes: windows of 1024, 2048 and 4096 rows)."""
    rng = np.random.default_rng(L1 + L2)
    a1, d1, c1, b1 = _chains(rng, nch1, L1, 0.6)
    a2, d2, c2, b2 = _chains(rng, nch2, L2, 0.6)
    a, d, c, b = (np.concatenate(q) for q in ((a1, a2), (d1, d2), (c1, c2), (b1, b2)))
    n1 = nch1 * L1
    ref = np.concatenate([_thomas64(a1, d1, c1, b1, L1), _thomas64(a2, d2, c2, b2, L2)])
    xe = _solve32(a, d, c, b, n1, L1, L2, 0)
    ee = np.abs(xe - ref).max() / np.abs(ref).max()
    assert ee < 1e-5
    for shape in (128, 256, 512):
        xw = _solve32(a, d, c, b, n1, L1, L2, shape)
        ew = np.abs(xw - ref).max() / np.abs(ref).max()
        print(f"  chains {nch1}x{L1} + {nch2}x{L2}: max error / max |x|: window ({shape} lanes) {ew:.2e}, "
              f"exact FP32 {ee:.2e}", flush=True)
        assert ew < 1e-5
        assert ew < 4 * ee + 1e-7


def test_window_shape_autotune_on_the_cavity():
    """The window-shape autotuner on the chains of a 512^2 cavity's last
    q-l iteration returns one of the three shapes and their times; the mixed
    run that continues with the chosen shape stays within 1e-6 of the FP64
    run (relative L2 per field)."""
    g, reg, prm = workloads.cavity(512, 1000.0, 2e-4, nPpeSolver=7, msorit=100, sortol=1e-10)
    z = np.zeros((g.ncol, g.ld))
    sol = api.Solver(g, reg, prm)
    sol.set_precision("mixed")
    sol.set_fields(z, z, z)
    sol.cold_start()
    sol.step(2)
    t = sol.trid_autotune(10)
    print(f"  window autotune at 512^2: {t}", flush=True)
    assert t["window_threads"] in (128, 256, 512)
    assert all(v > 0 for v in t["us"].values())
    sol.step(8)
    fm = sol.get_fields()
    sol.close()
    f64, _ = _run(g, reg, prm, "fp64", True, 10)
    S = np.s_[: g.ny + 2, : g.nx + 2]
    for name, a, b in zip("uvp", fm, f64):
        assert rel_l2(a[S], b[S]) < 1e-6, name


def test_window_solver_falls_back_on_weakly_dominant_chains():
    """Dominance ratio 0.95 (above the window solver's 0.9) in one window of a
    long chain: the bit-exact FP32 solver solves the whole batch, so x equals
    its result bit for bit."""
    rng = np.random.default_rng(95)
    a, d, c, b = _chains(rng, 1, 200000, 0.5)
    aw, dw, cw, bw = _chains(rng, 1, 3000, 0.95)
    a[100000:103000], d[100000:103000], c[100000:103000] = aw, dw, cw
    xw = _solve32(a, d, c, b, 0, 1, 200000, True)
    xe = _solve32(a, d, c, b, 0, 1, 200000, False)
    assert np.array_equal(xw, xe)
    # and with every window dominant the solvers differ (the window solver ran)
    a2, d2, c2, b2 = _chains(rng, 1, 200000, 0.5)
    assert not np.array_equal(_solve32(a2, d2, c2, b2, 0, 1, 200000, True),
                              _solve32(a2, d2, c2, b2, 0, 1, 200000, False))
