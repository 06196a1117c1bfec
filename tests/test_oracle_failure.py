"""Failure injection on the CPU: over-relaxation and non-finite input must end a solve with an
error status within a bounded number of outer iterations, never hang or run on with NaN.

The reference has no such check: a NaN relative norm fails every `<` test of its stop rule and the
loop keeps iterating to numOuterIter (src/Solver.cu:258-268).  The oracle and the device library
both raise IFL_ERR_DIVERGED from the FP64 residual reduction of the iteration whose residual is not
finite; tests/test_gpu_parity.py checks that the device stops at the same iteration.
"""
import numpy as np
import pytest

from incompflow import _abi
from oracle.oracle import Oracle
from tests.helpers import cavity, channel, cavity3d


def _over_relaxed(mk, which, alpha=1.9):
    c = mk()
    c.stop_rule = _abi.STOP_NONE
    if which in ("u", "both"):
        for q in range(3):
            c.relax[q] = alpha
    if which in ("p", "both"):
        c.relax[3] = alpha
    return c


CASES = {
    "cavity24": lambda: cavity(24),
    "channel32x12": lambda: channel(32, 12, 2000),
    "cavity3d": lambda: cavity3d(8, 8, 6),
}


@pytest.mark.parametrize("which", ["u", "p", "both"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_alpha_1_9_diverges_with_status(case, which):
    """alpha = 1.9 on the momentum or the p' relaxation makes SIMPLE blow up; the solve must stop
    with IFL_ERR_DIVERGED long before its 3000-iteration cap, and the last record it returns is the
    first non-finite (or overflowing) one."""
    cfg = _over_relaxed(CASES[case], which)
    o = Oracle(cfg)
    r = o.solve(3000, history=True)
    assert r["status"] == _abi.IFL_ERR_DIVERGED, r["status"]
    assert r["iterations"] < 200, r["iterations"]
    assert not r["converged"]
    h = r["history"]
    assert len(h) == r["iterations"]
    # every earlier record was finite: the stop came at the first bad residual
    assert all(np.isfinite(x.scaled) for x in h[:-1])


def test_stable_relaxation_does_not_trip_the_divergence_check():
    """The same cases at the reference's alpha 0.7 / 0.3 run their whole budget with status OK, so
    the check above is about the relaxation, not the case."""
    for mk in CASES.values():
        c = mk()
        c.stop_rule = _abi.STOP_NONE
        o = Oracle(c)
        r = o.solve(200, history=False)
        assert r["status"] == _abi.IFL_OK and r["iterations"] == 200


@pytest.mark.parametrize("field", [_abi.FIELD_U, _abi.FIELD_P])
def test_nan_in_initial_field_stops_at_first_iteration(field):
    c = cavity(16)
    o = Oracle(c)
    a = o.get(field)
    a.reshape(-1)[37] = np.nan
    o.set(field, a)
    r = o.solve(50, history=True)
    assert r["status"] == _abi.IFL_ERR_DIVERGED
    assert r["iterations"] == 1


def test_iteration_cap_without_convergence_reports_not_converged():
    """A cap far below what the case needs ends the solve with converged = False (the device
    returns IFL_ERR_NOT_CONVERGED for it when fail_if_not_converged is set)."""
    c = cavity(24)
    c.stop_rule = _abi.STOP_SCALED
    c.tol_residual = 1e-9
    o = Oracle(c)
    r = o.solve(30, history=False)
    assert r["iterations"] == 30 and not r["converged"] and r["status"] == _abi.IFL_OK
