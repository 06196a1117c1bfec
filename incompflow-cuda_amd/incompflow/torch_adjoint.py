"""torch.autograd binding of the steady solve and its discrete adjoint (thin test binding).

J = SteadyObjective.apply(profile, mu, backend) runs the forward solve to convergence with the
given inlet normal-velocity profile and viscosity and returns the objective; backward() runs the
adjoint of that solve and returns dJ/dprofile and dJ/dmu.  A backend provides

    solve(profile: np.ndarray, mu: float) -> run     # one forward solve, converged
    run.J                                            # objective at its converged state
    run.gradient() -> (np.ndarray, float)            # dJ/dprofile, dJ/dmu at that state

Each forward call returns its own run, kept by the autograd node, so a graph that evaluates the
objective several times (J(p1) + J(p2), difference quotients) takes every gradient at its own
converged state.  HipBackend is the library's (ifl_solve + ifl_adjoint through the C-ABI).
"""
import numpy as np
import torch

from . import _abi
from ._lib import Solver


class HipRun:
    """One converged forward solve on the device and, on demand, its adjoint."""

    def __init__(self, backend, profile, mu):
        self.backend = backend
        c = backend.cfg.copy()
        c.viscosity = float(mu)
        self.cfg = c
        self.solver = Solver(c)
        try:
            self.solver.set_face_profile(backend.face, np.asarray(profile, dtype=np.float64))
            self.forward_report = self.solver.solve(backend.max_iter, history=False)
            if self.forward_report["status"] != 0:
                raise RuntimeError("forward solve diverged")
            if backend.require_converged and not self.forward_report["converged"]:
                raise RuntimeError(
                    f"forward solve did not meet its stop rule in {self.forward_report['iterations']} outer "
                    "iterations: the adjoint needs a converged state (raise num_outer_iter or loosen the tolerance)")
            self.J = self._objective()
        except Exception:
            self.close()
            raise
        self.adjoint_report = None

    def _objective(self):
        s, c, be = self.solver, self.cfg, self.backend
        if be.objective == _abi.OBJ_PRESSURE_DROP:
            p = s.get(_abi.FIELD_P).astype(np.float64)
            return float(p[:, :, 0].mean() - p[:, :, c.nx - 1].mean())
        J = 0.0
        for fid, w in (be.weights or {}).items():
            J += float(np.dot(np.asarray(w, dtype=np.float64).ravel(), s.get(fid, shaped=False).astype(np.float64)))
        return J

    def gradient(self):
        if self.adjoint_report is None:
            be = self.backend
            if self.solver is None:
                raise RuntimeError("this forward solve has been closed")
            self.adjoint_report = self.solver.adjoint(be.objective, be.face, be.weights, tol=be.adjoint_tol,
                                                      grad_tol=be.adjoint_tol, check_every=500,
                                                      max_iter=be.adjoint_max_iter)
        r = self.adjoint_report
        return r["dJ_dprofile"], r["dJ_dmu"]

    def close(self):
        if self.solver is not None:
            self.solver.close()
            self.solver = None


class HipBackend:
    """Case config, inlet face, objective and tolerances; every solve() is a fresh device solver.

    require_converged=False accepts a forward solve that ran out of outer iterations without
    meeting its stop rule (e.g. a fixed iteration count, stop_rule = STOP_NONE); the adjoint is
    then taken at a state that is only as close to the fixed point as that solve got."""

    def __init__(self, cfg, face=_abi.WEST, objective=_abi.OBJ_PRESSURE_DROP, weights=None, adjoint_tol=1e-7,
                 max_iter=0, adjoint_max_iter=0, require_converged=True):
        self.cfg = cfg.copy()
        self.face = face
        self.objective = objective
        self.weights = weights
        self.adjoint_tol = adjoint_tol
        self.max_iter = max_iter
        self.adjoint_max_iter = adjoint_max_iter
        self.require_converged = require_converged

    def solve(self, profile, mu):
        return HipRun(self, profile, mu)


class SteadyObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, profile, mu, backend):
        run = backend.solve(profile.detach().to(torch.float64).cpu().numpy(), float(mu.detach()))
        ctx.run = run
        ctx.dtype = profile.dtype
        ctx.device = profile.device
        ctx.mu_dtype = mu.dtype
        ctx.mu_device = mu.device
        return torch.tensor(run.J, dtype=torch.float64, device=profile.device)

    @staticmethod
    def backward(ctx, gJ):
        gp, gmu = ctx.run.gradient()
        gJ = gJ.to(torch.float64)
        gprof = gJ.to(ctx.device) * torch.as_tensor(np.asarray(gp, dtype=np.float64), device=ctx.device)
        gm = gJ.to(ctx.mu_device) * torch.tensor(float(gmu), dtype=torch.float64, device=ctx.mu_device)
        return gprof.to(ctx.dtype), gm.to(ctx.mu_dtype), None


def steady_objective(profile, mu, backend):
    return SteadyObjective.apply(profile, mu, backend)
