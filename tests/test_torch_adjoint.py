"""The torch.autograd binding (incompflow/torch_adjoint.py) on the CPU, with an oracle backend:
autograd returns the backend's adjoint gradient, that gradient matches central differences of the
objective through the same binding, and every autograd node keeps the gradient of its own solve."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from incompflow import _abi
from incompflow.torch_adjoint import steady_objective
from test_oracle_adjoint import channel_case, converged, layout, pressure_drop


class OracleRun:
    """One FP64 oracle forward solve and its adjoint (the binding's per-call run protocol)."""

    def __init__(self, cfg, profile, mu):
        c = cfg.copy()
        c.viscosity = mu
        self.cfg = c
        self.o = converged(c, profile)
        self.J = pressure_drop(self.o, c)
        self._g = None

    def gradient(self):
        if self._g is None:
            self.o.adj_setup(0)
            for _ in range(200):
                if self.o.adj_iterate(250) < 1e-12:
                    break
            g = self.o.adj_gradient()
            _, _, poff = layout(self.cfg)
            self._g = (g[poff[_abi.WEST]:poff[_abi.WEST + 1]].copy(), float(g[-1]))
        return self._g


class OracleBackend:
    def __init__(self, cfg):
        self.cfg = cfg
        self.runs = []

    def solve(self, profile, mu):
        r = OracleRun(self.cfg, profile, mu)
        self.runs.append(r)
        return r


def test_autograd_gradient_is_the_adjoint_and_matches_differences():
    c, prof = channel_case(16, 8, noise=1e-2)
    be = OracleBackend(c)
    p = torch.tensor(prof, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(c.viscosity, dtype=torch.float64, requires_grad=True)
    J = steady_objective(p, mu, be)
    J.backward()
    gp, gmu = be.runs[-1].gradient()
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


def test_two_forward_calls_keep_their_own_gradients():
    """J(p1) + 2 J(p2) with one backward: each node's gradient is taken at its own converged state,
    not at the last solve's (the sum's gradient equals the gradients taken alone)."""
    c, prof = channel_case(16, 8, noise=1e-2)
    prof2 = prof * 1.5 + 0.002
    be = OracleBackend(c)
    p1 = torch.tensor(prof, dtype=torch.float64, requires_grad=True)
    p2 = torch.tensor(prof2, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(c.viscosity, dtype=torch.float64, requires_grad=True)
    J = steady_objective(p1, mu, be) + 2.0 * steady_objective(p2, mu, be)
    J.backward()
    g1, m1 = OracleRun(c, prof, c.viscosity).gradient()
    g2, m2 = OracleRun(c, prof2, c.viscosity).gradient()
    np.testing.assert_allclose(p1.grad.numpy(), g1, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(p2.grad.numpy(), 2.0 * g2, rtol=1e-12, atol=1e-15)
    assert mu.grad.item() == pytest.approx(m1 + 2.0 * m2, rel=1e-12)
    # the two states differ, so a gradient taken at the wrong one would be caught
    assert np.max(np.abs(g1 - g2)) > 1e-3 * np.max(np.abs(g1))
