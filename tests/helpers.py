"""Shared test helpers: golden fixtures -> case objects."""
from __future__ import annotations

import glob
import os

import numpy as np

from wolfd2_amd import case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def kernel_fixture_names():
    return sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "kernels_*.npz")))


def run_fixture_names():
    return sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "run_*.npz")))


def regions_from(fx: dict) -> case.Regions:
    nx, ny = int(fx["nx"]), int(fx["ny"])
    return case.Regions(nx=nx, ny=ny, nreg=fx["reg_nreg"], nRegBrd=np.asfortranarray(fx["reg_nRegBrd"]),
                        nRegType=np.asfortranarray(fx["reg_nRegType"]), nMomBdTp=np.asfortranarray(fx["reg_nMomBdTp"]),
                        dBCVal=np.asfortranarray(fx["reg_dBCVal"]), dPRporos=np.asfortranarray(fx["reg_dPRporos"]),
                        dPRporc1=np.asfortranarray(fx["reg_dPRporc1"]), dPRporc2=np.asfortranarray(fx["reg_dPRporc2"]))


def embed(a: np.ndarray, ld: int, ncol: int) -> np.ndarray:
    """Put a (ny+2, nx+2) fixture array into a zero (ncol, ld) field."""
    f = np.zeros((ncol, ld))
    f[: a.shape[0], : a.shape[1]] = a
    return f


def run_case(fx: dict, mnx: int | None = None):
    """Grid, regions and params of a whole-run fixture (deck of tests/golden/decks.py)."""
    x, y = fx["x"], fx["y"]
    ny, nx = x.shape
    deck = str(fx["deck"])
    g = case.compute_grid(x, y, mnx=mnx, mny=mnx)
    kw = {}
    for line in deck.splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "max_ql_iter":
            kw["mqiter"] = int(w[1])
        elif w[0] == "ql_tolerance":
            kw["qtol"] = float(w[1])
        elif w[0] == "max_sor_iter":
            kw["msorit"] = int(w[1])
        elif w[0] == "sor_tolerance":
            kw["sortol"] = float(w[1])
        elif w[0] == "sor_relaxation":
            kw["sorrel"] = float(w[1])
    cart = "cartesian_grid" in deck
    prm = case.Params.from_physical(float(fx["dt"]), float(fx["rho"]), float(fx["mu"]),
                                    nPpeSolver=int(fx["ppe_solver"]), cartesian=cart, **kw)
    walls, inlets, outlets = [], [], []
    for line in deck.splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "wall":
            tv = float(w[w.index("tangent_vel") + 1]) if "tangent_vel" in w else 0.0
            walls.append((int(w[1]), int(w[2]), w[3], "no_slip", tv, 0.0))
        elif w[0] == "inlet":
            nv = float(w[w.index("normal_vel") + 1]) if "normal_vel" in w else 0.0
            side = w[3]
            u, v = (nv, 0.0) if side in ("west", "east") else (0.0, nv)
            inlets.append((int(w[1]), int(w[2]), side, u, v, 0.0))
        elif w[0] == "outlet":
            kind = "fully_dev" if "fully_dev" in w else "mass_cons"
            pr = float(w[w.index("press") + 1]) if "press" in w else 0.0
            outlets.append((int(w[1]), int(w[2]), w[3], kind, 0.0, 0.0, pr))
    reg = case.make_regions(nx, ny, walls=walls, inlets=inlets, outlets=outlets)
    return g, reg, prm
