"""Throughput of the wolfd2 time step on one MI355X.

Workload: lid-driven square cavity, 1024x1024 cells (1025^2 nodes), Re = 1000,
dt = 2e-4, the reference's default tolerances (q-l 1e-4 / 20 iterations,
Poisson 1e-8 / 2000 sweeps) with red-black SOR, FP64, fluid at rest with the
lid at u = 1.  A step is one pass of main.f:690-972 over the whole grid:
nAuxMomentum (Douglas-Gunn ADI q-l iterations), the boundary fills, Ppe and
Project.  The timed region starts after a cold start and W warm-up steps,
with every field resident in HBM; it ends after K steps and a device sync.

Prints one JSON line (metric, value = Mcell-steps/s, roofline of the dominant
kernel, and the CPU baseline: the reference program compiled at -O3 on one
host core, timed on a one-step sample of the same cavity).
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024, help="cells per side")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
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
    u0, v0, p0 = sol.get_fields()
    if args.warmup:
        sol.step(args.warmup)
    t0 = time.perf_counter()
    tab = sol.step(args.steps)            # returns after a device synchronisation
    t1 = time.perf_counter()
    secs = t1 - t0
    cells = (nx - 1) * (ny - 1)
    value = cells * args.steps / secs / 1e6

    # dominant kernel: the red-black SOR kernel (k_sor_tiled, several sweeps per
    # launch on LDS tiles).  Mean launch duration from HIP events on the
    # context's stream over a second timed region of the same length.
    L = lib()
    import ctypes as C
    sweeps_per_launch = L.wd2_sor_sweeps_per_launch()
    avg_us, nl = C.c_double(0), C.c_long(0)
    L.wd2_ctx_kernel_timing(sol.ctx, 1, None, None)
    tab2 = sol.step(args.steps)
    L.wd2_ctx_kernel_timing(sol.ctx, 0, C.byref(avg_us), C.byref(nl))
    # algorithmic bytes per launch: a sweep reads p and the right-hand side and
    # writes p once per cell, 3 x 8 B per cell-sweep, times the sweeps per
    # launch (one colour kernel of the untiled path is half a sweep: 12 B)
    bytes_per_launch = (24.0 * sweeps_per_launch if sweeps_per_launch > 1 else 12.0) * cells
    achieved = bytes_per_launch / (avg_us.value * 1e-6) / 1e9 if avg_us.value > 0 else 0.0
    # roofline traffic from rocprofv3 counters (FETCH_SIZE + WRITE_SIZE per
    # dispatch, scaled by the copy-kernel calibration): see profiles/
    traffic = None
    prof = os.path.join(ROOT, "profiles", "sor_rb_1024_traffic.json")
    if os.path.exists(prof):
        try:
            traffic = json.load(open(prof)).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(nx, ny, g, reg, prm, u0, v0, p0)
    sol.close()

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * secs / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (cold-start lid-driven cavity)",
        "config": {"workload": "lid-driven cavity 1024x1024 cells, Re=1000, dt=2e-4, rb_sor, reference "
                               "tolerances (qtol 1e-4/20, sortol 1e-8/2000), FP64 bit-exact path",
                   "cells": cells, "path": "fp64",
                   "ql_iter_mean": float(np.mean(tab[:, 0])), "ppe_iter_mean": float(np.mean(tab[:, 1]))},
        "roofline": {"bound": "hbm", "kernel": (f"k_sor_tiled ({sweeps_per_launch} red-black SOR sweeps per launch)"
                                                if sweeps_per_launch > 1 else "k_sor_rb (one colour of a sweep)"),
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_us": round(avg_us.value, 3), "launches": int(nl.value),
                     "bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
