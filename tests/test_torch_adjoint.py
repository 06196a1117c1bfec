"""The torch.autograd binding (incompflow/torch_adjoint.py) on the CPU, with an oracle backend:
autograd returns the backend's adjoint gradient, and that gradient matches central differences
of the objective through the same binding."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from incompflow import _abi
from incompflow.torch_adjoint import steady_objective
from oracle.oracle import Oracle
from test_oracle_adjoint import channel_case, converged, layout, pressure_drop


class OracleBackend:
    """FP64 oracle forward solve and adjoint, behind the binding's backend protocol."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.o = None

    def solve(self, profile, mu):
        c = self.cfg.copy()
        c.viscosity = mu
        self.o = converged(c, profile)
        return pressure_drop(self.o, c)

    def gradient(self):
        self.o.adj_setup(0)
        for _ in range(200):
            if self.o.adj_iterate(250) < 1e-12:
                break
        g = self.o.adj_gradient()
        _, _, poff = layout(self.cfg)
        return g[poff[_abi.WEST]:poff[_abi.WEST + 1]].copy(), float(g[-1])


def test_autograd_gradient_is_the_adjoint_and_matches_differences():
    c, prof = channel_case(16, 8, noise=1e-2)
    be = OracleBackend(c)
    p = torch.tensor(prof, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(c.viscosity, dtype=torch.float64, requires_grad=True)
    J = steady_objective(p, mu, be)
    J.backward()
    gp, gmu = be.gradient()
    assert np.array_equal(p.grad.numpy(), gp)
    assert mu.grad.item() == gmu
    # central differences through the binding
    h = 1e-4 * c.viscosity
    with torch.no_grad():
        fd = (steady_objective(p, mu + h, be) - steady_objective(p, mu - h, be)).item() / (2 * h)
    assert abs(gmu - fd) <= 1e-5 * abs(fd)
    j = c.ny // 2
    hp = 1e-4 * 0.02
    e = torch.zeros_like(p)
    e[j] = hp
    with torch.no_grad():
        fdp = (steady_objective(p + e, mu, be) - steady_objective(p - e, mu, be)).item() / (2 * hp)
    assert abs(gp[j] - fdp) <= 1e-5 * np.abs(gp).max()
