"""Case 47 of the seeded 64-case batch (an 80x96 lid-driven cavity at Re 964) does not reach the
1e-4 scaled residual with the batch's settings, on the device and on the oracle alike: with alpha
0.7/0.3 and the reference's 20 p' sweeps per outer iteration (config.h:67-70) this SIMPLE iteration
settles into a limit cycle.  The case is steady and solvable: a stronger p' solve (40 sweeps) or a
smaller momentum relaxation (alpha_u 0.5) converges.  These tests pin that explanation on the FP32
oracle (the device equals the oracle bit for bit on this case, tests/test_gpu_batch.py)."""
import numpy as np

from incompflow import _abi
from oracle.oracle import Oracle


def _case47(**kw):
    c = _abi.uneven_batch_configs(tol_residual=1e-4, max_iter=20000)[47]
    assert (c.nx, c.ny) == (80, 96) and c.bc_type[_abi.WEST] == _abi.WALL
    assert abs(1.0 / c.viscosity - 964.1) < 0.1
    for k, v in kw.items():
        if k == "n_iter_p":
            c.n_iter[3] = v
        elif k == "alpha_u":
            for q in range(3):
                c.relax[q] = v
    return c


def test_batch_settings_give_a_limit_cycle():
    """3000 outer iterations: no convergence, and over the last 1500 the scaled residual keeps
    swinging between ~0.1 and ~0.6 instead of decaying (the device run stalls at 0.55 after 20000)."""
    c = _case47()
    c.stop_rule = _abi.STOP_NONE
    r = Oracle(c).solve(3000, history=True)
    s = np.array([x.scaled for x in r["history"]])
    tail = s[1500:]
    assert r["status"] == _abi.IFL_OK
    assert tail.min() > 0.05 and tail.max() > 0.3
    # no decay: the second half of the tail is as large as the first
    assert tail[750:].max() > 0.5 * tail[:750].max()


def test_stronger_p_solve_converges():
    r = Oracle(_case47(n_iter_p=40)).solve(3000, history=False)
    assert r["converged"] and r["iterations"] < 1000, r["iterations"]


def test_smaller_momentum_relaxation_converges():
    r = Oracle(_case47(alpha_u=0.5)).solve(3000, history=False)
    assert r["converged"] and r["iterations"] < 2000, r["iterations"]
