"""Autotuner: choices come from the candidate sets, the solver state after tuning is the state
before the trials (bit for bit), and the chosen settings are the ones the solve then uses."""
import numpy as np
import pytest

from helpers import STATE_FIELDS, assert_bitwise, cavity, cavity3d
from incompflow import Solver, _abi

pytestmark = pytest.mark.gpu


def _fields(s, cfg):
    return {f: s.get(f, shaped=False) for f in STATE_FIELDS if cfg.field_size(f) > 0}


def _case():
    c = cavity(48, 40, re=200.0)
    c.stop_rule = _abi.STOP_SCALED
    c.tol_residual = 1e-12
    return c


def test_autotune_choices_and_restored_state():
    c = _case()
    with Solver(c) as s:
        r = s.autotune(trial_iters=40)
        after = _fields(s, c)
        assert r["tile"] in (1, 2, 3, 4, 5)
        assert r["sweep_block"] in (1, 2, 4, 5)
        assert r["sweep_block_mom"] in (0, 4, 5) and r["sweep_block_mom"] in (0, r["sweep_block"])
        assert r["n_iter"][0] in (5, 10, 20) and r["n_iter"][1] == r["n_iter"][0]
        assert r["n_iter"][3] in (10, 20, 40)
        assert r["trials"] == 30 + 9 + 1
        assert r["us_per_iter"] > 0 and r["decades_per_s"] > 0
        rep = s.solve(30)
        assert rep["iterations"] == 30
        sweeps = np.array([[h.sweeps[q] for q in range(4)] for h in rep["history"]])
        assert sweeps[:, 3].max() <= r["n_iter"][3]
        assert sweeps[:, 0].max() <= r["n_iter"][0]
    # a fresh case had 20 warm-up iterations with its own settings and the stop test off
    w = c.copy()
    w.stop_rule = _abi.STOP_NONE
    with Solver(w) as s2:
        s2.solve(20, history=False)
        want = _fields(s2, w)
    for f in want:
        assert_bitwise(after[f], want[f], f"field {f}")


def test_autotune_launch_only_keeps_state_of_running_case():
    c = _case()
    with Solver(c) as s:
        s.solve(25, history=False)
        before = _fields(s, c)
        r = s.autotune(_abi.TUNE_LAUNCH)
        assert r["trials"] == 30
        assert r["n_iter"] == [10, 10, 10, 20]
        for f, v in _fields(s, c).items():
            assert_bitwise(v, before[f], f"field {f}")
        rep = s.solve(5)
        assert rep["iterations"] == 5


def test_autotune_launch_3d_picks_pcorr_sweeps_per_launch(tmp_path):
    """3D: the launch trials are one and two p' sweeps per launch; the state is kept and the
    solve continues with the choice, equal to a solve configured with it from the same state."""
    c = cavity3d(40, 20, 36)
    c.stop_rule = _abi.STOP_NONE
    path = str(tmp_path / "s.ifl")
    with Solver(c) as s:
        s.solve(25, history=False)
        before = _fields(s, c)
        s.save_state(path)
        r = s.autotune(_abi.TUNE_LAUNCH)
        assert r["trials"] == 2 and r["tile"] == 0 and r["sweep_block"] in (1, 2)
        for f, v in _fields(s, c).items():
            assert_bitwise(v, before[f], f"field {f}")
        s.solve(5, history=False)
        got = _fields(s, c)
    w = c.copy()
    w.sweep_block = r["sweep_block"]
    with Solver(w) as s2:
        s2.load_state(path)
        s2.solve(5, history=False)
        want = _fields(s2, w)
    for f in want:
        assert_bitwise(got[f], want[f], f"field {f}")


def test_converged_fields_do_not_depend_on_the_tuned_settings():
    """The fixed point solves M(uf)u = b(p), Rhie-Chow and div uf = 0 whatever the tile, sweeps per
    launch and sweeps per outer iteration: both solves to a 1e-5 scaled residual agree."""
    c = cavity(48, 40, re=200.0)
    c.stop_rule = _abi.STOP_SCALED
    c.tol_residual = 1e-5
    c.num_outer_iter = 100000
    with Solver(c) as a:
        ra = a.solve(0, history=False)
        fa = {f: a.get(f, shaped=False).astype(np.float64) for f in (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P)}
    with Solver(c) as b:
        r = b.autotune(trial_iters=60)
        rb = b.solve(0, history=False)
        fb = {f: b.get(f, shaped=False).astype(np.float64) for f in (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P)}
    assert ra["converged"] and rb["converged"], (ra["iterations"], rb["iterations"], r)
    fa[_abi.FIELD_P] -= fa[_abi.FIELD_P].mean()
    fb[_abi.FIELD_P] -= fb[_abi.FIELD_P].mean()
    for f in fa:
        err = np.abs(fa[f] - fb[f]).max() / np.abs(fa[f]).max()
        assert err < 1e-4, (f, err, r)
