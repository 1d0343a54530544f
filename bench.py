"""Throughput of the wolfd2 time step on one MI355X: mixed precision against FP64.

Workload: lid-driven square cavity, 1024x1024 cells (1025^2 nodes), Re = 1000,
dt = 2e-4, the reference's default tolerances (q-l 1e-4 / 20 iterations,
Poisson 1e-8 / 2000 sweeps) with red-black SOR, fluid at rest with the lid at
u = 1.  A step is one pass of main.f:690-972 over the whole grid:
nAuxMomentum (Douglas-Gunn ADI q-l iterations), the boundary fills, Ppe and
Project.

Both arithmetic paths start from the same cold-start state: the FP64 path
(the reference's arithmetic, bit for bit) and the mixed path (FP64 residual
and pressure accumulation, FP32 red-black SOR sweeps).  Each timed region
starts after W warm-up steps with every field resident in HBM and ends after
K steps and a device sync.  value = the mixed path's Mcell-steps/s; the FP64
path's throughput and the per-field relative L2 / max error of the mixed
path's u, v, p against the FP64 path's after the same steps are reported
beside it.

Prints one JSON line: metric, value, roofline of the dominant kernel (the
mixed path's tiled SOR kernel; the FP64 one under fp64_roofline), and the CPU
baseline: the reference program compiled at -O3 on one host core, timed on a
one-step sample of the same cavity.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from wolfd2_amd import api, case  # noqa: E402
from wolfd2_amd._lib import lib  # noqa: E402

METRIC = "Mcell-steps/s @1024^2: mixed-precision vs FP64 GPU at 1e-6 L2; % HBM peak"
HBM_PEAK_GBS = 8000.0


def host_cpu() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(nx: int, ny: int, g, reg, prm, u0, v0, p0, timeout: float = 600.0) -> dict | None:
    """The reference program (oracle/_ref/fast/wolfd2, -O3 -ffast-math) on one
    host core: T(1 step) - T(0 steps), both started from the same restart file
    (the cavity after its cold start), so the difference is one time step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    exe = os.path.join(ROOT, "oracle", "_ref", "fast", "wolfd2")
    if not os.path.exists(exe):
        return None
    import refdecks
    cells = (nx - 1) * (ny - 1)
    x, y = case.uniform_nodes(nx, ny)
    times = {}
    with tempfile.TemporaryDirectory(prefix="wd2_cpu_") as rd:
        fields = {"u": u0[: ny + 2, : nx + 2], "v": v0[: ny + 2, : nx + 2], "p": p0[: ny + 2, : nx + 2]}
        refdecks.write_restart(os.path.join(rd, "cold.rst"), 0, 0.0, nx, ny, fields)
        for nts in (0, 1):
            d = refdecks.Deck(nts=nts, dt=2e-4, mu=1e-3, ppe_solver="rb_sor", restart_in="cold.rst",
                              print_diff_freq=1)
            d.write(rd, x, y)
            t0 = time.perf_counter()
            r = subprocess.run(["taskset", "-c", "0", exe], cwd=rd, capture_output=True, text=True, timeout=timeout)
            times[nts] = time.perf_counter() - t0
            if r.returncode != 0:
                return None
            if nts == 1:
                tab = refdecks.parse_print_diff(r.stdout)
    dt = times[1] - times[0]
    return {"value": round(cells / dt / 1e6, 6), "unit": "Mcell-steps/s", "cores": 1, "kind": "reference",
            "sample": (f"one time step of the same 1024^2 cavity from its cold-start state (restart file), "
                       f"reference sources built -O3 -ffast-math, taskset 1 core of '{host_cpu()}'; "
                       f"T(1 step)-T(0 steps) = {times[1]:.1f}-{times[0]:.1f} s; "
                       f"QlIt={int(tab[0, 2]) if len(tab) else -1}, PpeIt={int(tab[0, 3]) if len(tab) else -1}")}


def copy_calibration(mib: int) -> int:
    import torch
    n = mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst, src))
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 5, "ok": ok}))
    return 0 if ok else 1


def _traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 counters (FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        return json.load(open(path)).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def run_path(sol, precision: str, fields0, warmup: int, steps: int, cells: int) -> dict:
    """W warm-up steps and K timed steps of one arithmetic path from the same
    state, then K more steps with the dominant kernel's launches timed by HIP
    events on the context's stream."""
    import ctypes as C
    L = lib()
    sol.set_precision(precision)
    sol.set_fields(*fields0)
    if warmup:
        sol.step(warmup)
    t0 = time.perf_counter()
    tab = sol.step(steps)                 # returns after a device synchronisation
    secs = time.perf_counter() - t0
    fields = sol.get_fields()
    avg_us, nl = C.c_double(0), C.c_long(0)
    L.wd2_ctx_kernel_timing(sol.ctx, 1, None, None)
    sol.step(steps)
    L.wd2_ctx_kernel_timing(sol.ctx, 0, C.byref(avg_us), C.byref(nl))
    if precision == "mixed":
        # one launch: FP64 defect correction (read p and b, write p: 24 B per
        # cell) around k FP32 sweeps of the correction (read e and r, write e:
        # 12 B per cell and sweep)
        k = L.wd2_sor_mixed_sweeps_per_launch()
        bpl = (24.0 + 12.0 * k) * cells
        kern = f"k_sor_tiled_mx ({k} FP32 red-black SOR sweeps per launch on an FP64 residual)"
        traffic = _traffic("k_sor_tiled_mx_cavity1024_traffic.json")
    else:
        # a sweep reads p and the right-hand side and writes p: 24 B per cell
        k = L.wd2_sor_sweeps_per_launch()
        bpl = (24.0 * k if k > 1 else 12.0) * cells
        kern = (f"k_sor_tiled ({k} red-black SOR sweeps per launch)" if k > 1
                else "k_sor_rb (one colour of a sweep)")
        traffic = _traffic("k_sor_tiled_cavity1024_traffic.json")
    achieved = bpl / (avg_us.value * 1e-6) / 1e9 if avg_us.value > 0 else 0.0
    return {"value": cells * steps / secs / 1e6, "ms_per_step": 1e3 * secs / steps, "fields": fields,
            "ql": float(np.mean(tab[:, 0])), "ppe": float(np.mean(tab[:, 1])),
            "roofline": {"bound": "hbm", "kernel": kern, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                         "avg_launch_us": round(avg_us.value, 3), "launches": int(nl.value),
                         "bytes_per_launch": bpl}}


def dropin_rate(g, reg, prm, fields0, steps: int = 2) -> dict:
    """Time steps of main.f:690-972 driven from the host through the drop-in
    entry points (the reference's own call sequence on host arrays, so every
    call copies its inputs to the GPU and its outputs back): the
    PCIe-inclusive rate of an unmodified Fortran driver, FP64."""
    nx, ny = g.nx, g.ny
    u, v, p = (a.copy() for a in fields0)
    d = np.zeros_like(u)

    def step():
        nonlocal u, v
        pn, un, vn, dn = p.copy(), u.copy(), v.copy(), d.copy()
        us, vs = u.copy(), v.copy()
        api.naux_momentum(reg, g, prm, d, dn, un, vn, us, vs)
        api.vel_bound_cond(reg, us, vs)
        api.pres_bound_cond(reg, p)
        api.ppe(reg, g, prm, us, vs, p)
        api.pres_bound_cond(reg, p)
        api.project(reg, g, prm.dk, p, us, vs)
        api.vel_bound_cond(reg, us, vs)
        api.pres_bound_cond(reg, p)
        u, v = us, vs
        api.vel_bound_cond(reg, u, v)
        api.pres_bound_cond(reg, p)
        for a, b in ((pn, p), (un, u), (vn, v)):
            api.diff_max_norm(nx, ny, a, b)

    step()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    secs = time.perf_counter() - t0
    cells = (nx - 1) * (ny - 1)
    return {"value": round(cells * steps / secs / 1e6, 4), "unit": "Mcell-steps/s",
            "ms_per_step": round(1e3 * secs / steps, 3), "steps": steps, "path": "fp64 drop-in, host arrays"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024, help="cells per side")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (the reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (the hot path is not sharded); "
                                   "run with --gpus 1", "n_gpus": args.gpus}))
        return 2

    lib().wd2_set_verbose(0)      # stdout carries the JSON line only
    nx = ny = args.n + 1
    x, y = case.uniform_nodes(nx, ny)
    g = case.compute_grid(x, y)
    reg = case.cavity_regions(nx, ny)
    prm = case.Params.from_physical(2e-4, 1.0, 1e-3, nPpeSolver=5)
    sol = api.Solver(g, reg, prm)
    z = np.zeros((g.ncol, g.ld))
    sol.set_fields(z, z, z)
    sol.cold_start()
    fields0 = sol.get_fields()
    cells = (nx - 1) * (ny - 1)
    res = {prec: run_path(sol, prec, fields0, args.warmup, args.steps, cells) for prec in ("fp64", "mixed")}
    mx, f64 = res["mixed"], res["fp64"]
    S = np.s_[: ny + 2, : nx + 2]
    err = {}
    for name, a, b in zip("uvp", mx["fields"], f64["fields"]):
        a, b = a[S], b[S]
        nb = float(np.linalg.norm(b))
        err[name] = {"rel_l2": float(np.linalg.norm(a - b)) / (nb if nb > 0 else 1.0),
                     "max_abs": float(np.max(np.abs(a - b)))}
    sol.close()
    dropin = dropin_rate(g, reg, prm, fields0)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(nx, ny, g, reg, prm, *fields0)

    out = {
        "metric": METRIC,
        "value": round(mx["value"], 4),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mx["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (cold-start lid-driven cavity)",
        "config": {"workload": "lid-driven cavity 1024x1024 cells, Re=1000, dt=2e-4, rb_sor, reference "
                               "tolerances (qtol 1e-4/20, sortol 1e-8/2000); mixed path: FP64 fields, FP64 "
                               "Poisson residual and pressure accumulation, FP32 red-black SOR sweeps, FP32 "
                               "momentum stencils and ADI chain solves",
                   "cells": cells, "path": "mixed",
                   "fp64_value": round(f64["value"], 4), "fp64_ms_per_step": round(f64["ms_per_step"], 4),
                   "speedup_vs_fp64": round(mx["value"] / f64["value"], 3),
                   "error_vs_fp64": err,
                   "within_1e-6_l2": all(e["rel_l2"] <= 1e-6 for e in err.values()),
                   "ql_iter_mean": mx["ql"], "ppe_iter_mean": mx["ppe"],
                   "fp64_ql_iter_mean": f64["ql"], "fp64_ppe_iter_mean": f64["ppe"]},
        "roofline": mx["roofline"],
        "fp64_roofline": f64["roofline"],
        "dropin_pcie_inclusive": dropin,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = round(mx["value"] / cpu["value"], 1)
        out["fp64_gpu_over_cpu"] = round(f64["value"] / cpu["value"], 1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
