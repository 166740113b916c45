"""Sensitivity of the mean hot-wall Nusselt number to the hot-wall
temperature profile, through the discrete adjoint of the energy steps.

``HotWallProblem`` holds one cavity on the GPU and a fixed start state.  A
forward evaluation sets the hot-wall (side a) Dirichlet values from the
control vector (piecewise-linear in height), restarts from the start state,
runs K single-pass SIMPLE iterations with checkpoints (ixh_adjoint_record)
and returns the mean Nusselt number on the hot wall (ixh_nusselt_x).  The
backward pass runs ixh_adjoint_energy with dJ/dT = the Nusselt integral's
weights and maps the boundary-node adjoints back to the controls.
``HotWallNusselt`` exposes this as a torch.autograd.Function.

By default the backward pass is the adjoint of the whole step
(ixh_adjoint_coupled): the exact gradient of the discrete K-step map, buoyancy
included.  With ``coupled=False`` it runs through the energy steps only with
the velocities held fixed (ixh_adjoint_energy): exact when the flow does not
depend on T, the frozen-flow part of the gradient otherwise.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from .case import Case
from .simple import Cavity

BC_DIRI = 1


def nusselt_x_weights(mesh, wall: str = "west") -> np.ndarray:
    """W with Nu = sum(W * T) for the vertical-wall mean Nusselt number of
    ixh_nusselt_x (quadratic through the three nodes nearest the wall,
    evaluated at the wall, weighted by dyp and divided by the height)."""
    mi, nj = mesh.mi, mesh.nj
    xp, xu, yv = mesh.xp, mesh.xu, mesh.yv
    W = np.zeros((mi + 1, nj + 1), order="F")
    hgt = yv[nj - 1] - yv[0]
    if wall == "west":
        i1, xw = 0, xu[0]
    else:
        i1, xw = mi - 2, xu[mi - 1]
    x1, x2, x3 = xp[i1], xp[i1 + 1], xp[i1 + 2]
    h1, h2, L = x2 - x1, x3 - x2, x3 - x1
    w = 2.0 * xw - x1 - x2
    for jj in range(2, nj + 1):           # 1-based rows 2..nj
        dyp = yv[jj - 1] - yv[jj - 2]
        c1 = dyp * (1.0 - w / L) * (-1.0 / h1)
        c2 = dyp * ((1.0 - w / L) / h1 - (w / L) / h2)
        c3 = dyp * (w / L) / h2
        W[i1, jj - 1] += -c1 / hgt
        W[i1 + 1, jj - 1] += -c2 / hgt
        W[i1 + 2, jj - 1] += -c3 / hgt
    return W


def control_matrix(mesh, ncontrols: int) -> np.ndarray:
    """B (nj+1, ncontrols): hot-wall node values from control values placed
    uniformly over the height, linear in between, constant beyond the ends
    (the ghost nodes)."""
    y = np.asarray(mesh.yp, dtype=np.float64)
    yc = np.linspace(mesh.yv[0], mesh.yv[-1], ncontrols)
    B = np.zeros((len(y), ncontrols))
    for r, yy in enumerate(y):
        if yy <= yc[0]:
            B[r, 0] = 1.0
        elif yy >= yc[-1]:
            B[r, -1] = 1.0
        else:
            k = int(np.searchsorted(yc, yy) - 1)
            t = (yy - yc[k]) / (yc[k + 1] - yc[k])
            B[r, k] = 1.0 - t
            B[r, k + 1] = t
    return B


class HotWallProblem:
    """J(c) = mean hot-wall Nusselt number after ``steps`` SIMPLE iterations
    from a fixed start, with hot-wall temperatures B c."""

    def __init__(self, case: Case, ncontrols: int = 64, steps: int = 400, every: int = 20,
                 solver: str = "pcr", start: Optional[Dict[str, np.ndarray]] = None, coupled: bool = True):
        if case.params.iter_ecuaci_max >= 0 or case.params.iter_simple_max >= 0:
            raise ValueError("the adjoint needs single-pass loops")
        self.case = case
        self.steps, self.every = steps, every
        self.coupled = coupled
        self.cav = Cavity(case, solver=solver)
        self.B = control_matrix(case.mesh, ncontrols)
        self.W = nusselt_x_weights(case.mesh)
        typ, _ = case.bcs["t"].expand(case.mesh)["a"]
        self.types_a = np.asarray(typ, dtype=np.int32)
        if np.any(self.types_a[1:-1] != BC_DIRI):
            raise ValueError("the hot wall (side a) must be Dirichlet for T")
        self.start = start if start is not None else {f: self.cav.get(f) for f in ("u", "v", "t", "p")}
        self._recorded: Optional[np.ndarray] = None  # controls of the run the cavity holds

    def wall_values(self, controls) -> np.ndarray:
        return self.B @ np.asarray(controls, dtype=np.float64)

    def forward(self, controls) -> float:
        c = self.cav
        controls = np.array(controls, dtype=np.float64)
        self._recorded = None
        c.set_bc("t", "a", self.types_a, self.wall_values(controls))
        for f, arr in self.start.items():
            c.set(f, arr)
            if f in ("u", "v", "t"):
                c.set(f + "_ant", arr)
        c.adjoint_record(self.steps, self.every)
        self._recorded = controls
        return c.nusselt_x()[0]

    def gradient(self, scale: float = 1.0, controls=None) -> np.ndarray:
        """dJ/dc times ``scale`` at ``controls``, or at the controls of the
        last forward evaluation when None.  The adjoint runs backward over
        the run the cavity recorded last, so when that run belongs to other
        controls (another forward evaluation came in between) the run is
        recorded again at ``controls`` first."""
        if controls is not None:
            controls = np.asarray(controls, dtype=np.float64)
            if self._recorded is None or not np.array_equal(controls, self._recorded):
                self.forward(controls)
        elif self._recorded is None:
            raise RuntimeError("gradient() needs a forward evaluation first")
        g, _ = self.cav.adjoint_energy(scale * self.W, coupled=self.coupled)
        return self.B.T @ g["a"]


try:
    import torch

    class HotWallNusselt(torch.autograd.Function):
        """Nu = HotWallNusselt.apply(controls, problem); backward through the
        discrete adjoint on the GPU."""

        @staticmethod
        def forward(ctx, controls, problem: HotWallProblem):
            c = controls.detach().cpu().double().numpy().copy()
            nu = problem.forward(c)
            ctx.problem = problem
            ctx.controls = c
            return controls.new_tensor(nu, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            # at this evaluation's own controls, even when other forward
            # evaluations of the same problem ran since
            g = ctx.problem.gradient(float(grad_out), controls=ctx.controls)
            return torch.from_numpy(g).to(device=grad_out.device, dtype=torch.float64), None
except ImportError:  # torch is optional for the numpy interface
    pass
