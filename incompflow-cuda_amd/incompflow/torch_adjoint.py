"""torch.autograd binding of the steady solve and its discrete adjoint (thin test binding).

J = SteadyObjective.apply(profile, mu, backend) runs the forward solve to convergence with the
given inlet normal-velocity profile and viscosity and returns the objective; backward() runs the
adjoint once and returns dJ/dprofile and dJ/dmu.  A backend provides

    solve(profile: np.ndarray, mu: float) -> float     # J at the converged state
    gradient() -> (np.ndarray, float)                  # dJ/dprofile, dJ/dmu at that state

HipBackend is the library's (ifl_solve + ifl_adjoint through the C-ABI).
"""
import numpy as np
import torch

from . import _abi
from ._lib import Solver


class HipBackend:
    """A fresh solver per forward call: case config, inlet face, objective and tolerances."""

    def __init__(self, cfg, face=_abi.WEST, objective=_abi.OBJ_PRESSURE_DROP, weights=None, adjoint_tol=1e-7,
                 max_iter=0, adjoint_max_iter=0):
        self.cfg = cfg.copy()
        self.face = face
        self.objective = objective
        self.weights = weights
        self.adjoint_tol = adjoint_tol
        self.max_iter = max_iter
        self.adjoint_max_iter = adjoint_max_iter
        self.solver = None
        self.forward_report = None
        self.adjoint_report = None

    def solve(self, profile, mu):
        if self.solver is not None:
            self.solver.close()
        c = self.cfg.copy()
        c.viscosity = float(mu)
        self.solver = Solver(c)
        self.solver.set_face_profile(self.face, np.asarray(profile, dtype=np.float64))
        self.forward_report = self.solver.solve(self.max_iter, history=False)
        if self.forward_report["status"] != 0:
            raise RuntimeError("forward solve diverged")
        return self.objective_value()

    def objective_value(self):
        s, c = self.solver, self.cfg
        if self.objective == _abi.OBJ_PRESSURE_DROP:
            p = s.get(_abi.FIELD_P).astype(np.float64)
            return float(p[:, :, 0].mean() - p[:, :, c.nx - 1].mean())
        J = 0.0
        for fid, w in (self.weights or {}).items():
            J += float(np.dot(np.asarray(w, dtype=np.float64).ravel(), s.get(fid, shaped=False).astype(np.float64)))
        return J

    def gradient(self):
        r = self.solver.adjoint(self.objective, self.face, self.weights, tol=self.adjoint_tol,
                                grad_tol=self.adjoint_tol, check_every=500, max_iter=self.adjoint_max_iter)
        self.adjoint_report = r
        return r["dJ_dprofile"], r["dJ_dmu"]


class SteadyObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, profile, mu, backend):
        J = backend.solve(profile.detach().to(torch.float64).cpu().numpy(), float(mu.detach()))
        ctx.backend = backend
        ctx.dtype = profile.dtype
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        gp, gmu = ctx.backend.gradient()
        gJ = gJ.to(torch.float64)
        return (gJ * torch.from_numpy(np.asarray(gp, dtype=np.float64))).to(ctx.dtype), gJ * gmu, None


def steady_objective(profile, mu, backend):
    return SteadyObjective.apply(profile, mu, backend)
