"""torch.autograd harness over the HIP path's adjoint (wd2_adj_*).

``CavityRun`` runs nsteps time steps of a one-region cavity from fixed base
fields with two controls: the lid values (the north side's u profile on
lines i = 2..nx) and the initial velocity on the points a step updates (u on
i = 2..nx-1, j = 2..ny; v on i = 2..nx, j = 2..ny-1).  ``LidRun`` is the
torch.autograd.Function: its forward runs the steps on the GPU with
checkpoints and returns the final u, v on the same points; its backward runs
the adjoint pass.  Tensors are float64 on the CPU (the fields stay on the GPU
between the passes; only the controls and outputs cross to the host).
"""
from __future__ import annotations

import numpy as np
import torch

from . import api, case, workloads


class CavityRun:
    def __init__(self, g, reg, prm, nsteps: int, checkpoint_every: int = 20, base=None, lid_base=None):
        self.g, self.reg, self.prm, self.nsteps = g, reg, prm, nsteps
        self.solver = api.Solver(g, reg, prm)
        self.adj = api.Adjoint(self.solver, checkpoint_every)
        z = np.zeros((g.ncol, g.ld))
        self.base = tuple(np.array(f, dtype=np.float64) for f in (base if base is not None else (z, z, z)))
        self.lid_base = np.full(g.nx + 2, reg.dBCVal[0, 0, case.NORTH - 1, case.VAR_U - 1]) \
            if lid_base is None else np.array(lid_base, dtype=np.float64)
        self.su, self.sv = workloads.interior_u(g), workloads.interior_v(g)
        self.nlid = g.nx - 1

    def close(self):
        self.adj.close()
        self.solver.close()

    def _setup(self, lid, u0, v0):
        g = self.g
        prof = self.lid_base.copy()
        prof[2:g.nx + 1] = lid
        self.solver.set_bc_profile(1, 1, case.NORTH, case.VAR_U, prof)
        u, v, p = (f.copy() for f in self.base)
        u[self.su] = u0
        v[self.sv] = v0
        self.solver.set_fields(u, v, p)

    def run(self, lid, u0, v0, checkpoints: bool = True):
        """Final (u, v) on the updated points; with checkpoints, ready for grad."""
        self._setup(lid, u0, v0)
        if checkpoints:
            self.tab = self.adj.forward(self.nsteps)
        else:
            self.tab = self.solver.step(self.nsteps)[:, :2].astype(np.int32)
        u, v, _ = self.solver.get_fields()
        return u[self.su].copy(), v[self.sv].copy()

    def grad(self, gu, gv):
        """Adjoints of (lid, u0, v0) given those of the last run's outputs."""
        g = self.g
        ub, vb = np.zeros((g.ncol, g.ld)), np.zeros((g.ncol, g.ld))
        ub[self.su] = gu
        vb[self.sv] = gv
        r = self.adj.backward(ub, vb)
        lid_bar = r["bc_bar"][case.NORTH - 1, case.VAR_U - 1, 2:g.nx + 1].copy()
        return lid_bar, r["u0_bar"][self.su].copy(), r["v0_bar"][self.sv].copy(), r["stats"]


class LidRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lid, u0, v0, run: CavityRun):
        uN, vN = run.run(lid.detach().cpu().numpy(), u0.detach().cpu().numpy(), v0.detach().cpu().numpy())
        ctx.run = run
        return torch.from_numpy(uN), torch.from_numpy(vN)

    @staticmethod
    def backward(ctx, gu, gv):
        run = ctx.run
        g = run.g
        gu = np.zeros((g.ny - 1, g.nx - 2)) if gu is None else gu.detach().cpu().numpy()
        gv = np.zeros((g.ny - 2, g.nx - 1)) if gv is None else gv.detach().cpu().numpy()
        lb, ub, vb, stats = run.grad(gu, gv)
        run.last_stats = stats
        return torch.from_numpy(lb), torch.from_numpy(ub), torch.from_numpy(vb), None
