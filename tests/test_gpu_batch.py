"""Batched solves: every case of a batch must equal its own single-case solve bit for bit
(fields, iteration count, convergence flag), whatever the other cases in the batch are."""
import sys

import numpy as np
import pytest

from helpers import STATE_FIELDS, assert_bitwise, cavity, cavity3d, channel, duct3d
from incompflow import Batch, Solver, _abi
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fields(s, cfg):
    return {f: s.get(f, shaped=False) for f in STATE_FIELDS if cfg.field_size(f) > 0}


def _check_batch(cfgs, max_iter, label):
    _log(f"[batch] {label}: {len(cfgs)} cases, max_iter={max_iter}")
    with Batch(cfgs) as b:
        st, reps = b.solve(max_iter)
        got = [_fields(b.cases[i], cfgs[i]) for i in range(len(cfgs))]
    for i, c in enumerate(cfgs):
        with Solver(c) as s:
            r = s.solve(max_iter, history=False)
            want = _fields(s, c)
        assert reps[i]["iterations"] == r["iterations"], (label, i)
        assert reps[i]["converged"] == r["converged"], (label, i)
        assert reps[i]["final_scaled"] == r["final_scaled"], (label, i)
        for f in want:
            assert_bitwise(got[i][f], want[f], f"{label} case {i} field {f}")
    return st, reps


def _mixed_2d(tile=3):
    cs = [channel(24, 12, 30), cavity(20, 16), cavity(70, 37, re=400.0), cavity(33, 18, re=20.0),
          channel(64, 16, 30), cavity(64, 64, re=1000.0)]
    cs[4].relax[0] = cs[4].relax[1] = 0.6
    cs[5].relax[3] = 0.2
    for c in cs:
        c.tile = tile
        c.num_outer_iter = 30
    return cs


@pytest.mark.parametrize("tile", [1, 3, 4, 5])
def test_batch_fixed_iterations_equal_single_solves(tile):
    _check_batch(_mixed_2d(tile), 30, f"fixed tile={tile}")


def test_batch_128x32_tiles_five_sweep_blocks():
    """1024-thread 128x32 tiles in a batch of small uneven cases (mostly partial tiles), five p' and
    five momentum sweeps per launch (odd launch counts: 10/7 sweeps of u/v)."""
    cs = _mixed_2d(5)
    for c in cs:
        c.sweep_block, c.sweep_block_mom = 5, 5
        c.n_iter[1] = 7
        c.num_outer_iter = 20
    _check_batch(cs, 20, "128x32 K=5 Km=5")


def test_batch_stop_rules_caps_and_repacking():
    """Cases stop at different iterations (scaled residual, own iteration caps); the batch
    repacks its graph as they stop (check every 8 iterations)."""
    cs = []
    for k, (n, m, re, tol, cap) in enumerate([(24, 24, 10.0, 1e-2, 400), (40, 20, 100.0, 1e-2, 400),
                                               (32, 32, 400.0, 3e-3, 400), (20, 20, 50.0, 1e-3, 9),
                                               (48, 24, 50.0, 1e-9, 17), (36, 36, 100.0, 1e-9, 60)]):
        c = cavity(n, m, re=re)
        c.stop_rule = _abi.STOP_SCALED
        c.tol_residual = tol
        c.num_outer_iter = cap
        c.check_every = 8
        c.tile = 2
        cs.append(c)
    faithful = channel(30, 10, 400)
    faithful.stop_rule = _abi.STOP_FAITHFUL
    faithful.check_every = 8
    faithful.tile = 2
    cs.append(faithful)
    st, reps = _check_batch(cs, 0, "stop rules")
    assert st == 0
    its = [r["iterations"] for r in reps]
    assert len(set(its)) >= 4, its          # the cases really stop at different iterations
    assert its[3] == 9 and its[4] == 17 and its[5] == 60


def test_batch_case_matches_oracle():
    """Tie the batched path to the oracle directly on one case of a mixed batch."""
    cs = _mixed_2d(1)[:3]
    for c in cs:
        c.num_outer_iter = 12
    with Batch(cs) as b:
        b.solve(12)
        got = _fields(b.cases[2], cs[2])
    o = Oracle(cs[2])
    o.solve(12)
    for f in got:
        assert_bitwise(got[f], o.get(f).ravel(), f"oracle field {f}")


def test_batch_3d():
    cs = [cavity3d(10, 8, 6), duct3d(12, 6, 5), cavity3d(20, 17, 33, re=300.0)]
    for c in cs:
        c.num_outer_iter = 8
    _check_batch(cs, 8, "3d")


def test_batch_3d_two_pcorr_sweeps_per_launch():
    """sweep_block 2 in 3D: the batched two-sweep p' launches (several columns and chunks per
    case, 7 p' sweeps: three two-sweep launches and a one-sweep launch) equal each case's solve."""
    cs = [cavity3d(40, 20, 36), duct3d(35, 9, 70), cavity3d(20, 17, 33, re=300.0)]
    for c in cs:
        c.num_outer_iter = 6
        c.sweep_block = 2
        c.n_iter[3] = 7
    _check_batch(cs, 6, "3d p' block 2")


def test_batch_jacobi():
    cs = _mixed_2d(1)[:4]
    for c in cs:
        c.smoother = _abi.SMOOTHER_JACOBI
        c.num_outer_iter = 10
    _check_batch(cs, 10, "jacobi")


@pytest.mark.parametrize("sweep_block,p_refine", [(2, 0), (1, 1), (2, 1), (4, 1), (5, 1)])
def test_batch_sweep_block_and_refine(sweep_block, p_refine):
    cs = _mixed_2d(2)
    for c in cs:
        c.sweep_block = sweep_block
        c.p_refine = p_refine
        c.n_iter[1] = 7           # u and v with different sweep counts
        c.num_outer_iter = 15
    _check_batch(cs, 15, f"K={sweep_block} refine={p_refine}")


def test_batch_inlet_profile_and_reuse():
    """Face profiles set on a case handle reach the batch; a second solve continues."""
    cs = [channel(24, 12, 10), channel(40, 16, 10)]
    for c in cs:
        c.tile = 4
    prof = [0.001 * (1.0 + 0.1 * np.sin(np.arange(c.ny))) for c in cs]
    with Batch(cs) as b:
        for i in range(2):
            b.cases[i].set_face_profile(_abi.WEST, prof[i])
        b.solve(6)
        b.solve(4)
        got = [_fields(b.cases[i], cs[i]) for i in range(2)]
    for i, c in enumerate(cs):
        with Solver(c) as s:
            s.set_face_profile(_abi.WEST, prof[i])
            s.solve(6, history=False)
            s.solve(4, history=False)
            want = _fields(s, c)
        for f in want:
            assert_bitwise(got[i][f], want[f], f"profile case {i} field {f}")


def _uneven(k, max_iter):
    cs = _abi.uneven_batch_configs(tol_residual=1e-4, max_iter=max_iter)[:k]
    for c in cs:
        c.sweep_block = 4
    return cs


def test_uneven_batch_subset_equals_single_solves_and_oracle():
    """The first eight cases of the seeded uneven batch (cavities and channels, 64..512 cells a
    side, Re 10..1000), 200 iterations each: every case equals its single solve, and the smallest
    one equals the oracle."""
    cs = _uneven(8, 200)
    _check_batch(cs, 200, "uneven subset")
    i = min(range(len(cs)), key=lambda q: cs[q].nx * cs[q].ny)
    with Batch(cs) as b:
        b.solve(200)
        got = _fields(b.cases[i], cs[i])
    o = Oracle(cs[i])
    o.solve(200, history=False)
    for f in got:
        assert_bitwise(got[f], o.get(f).ravel(), f"uneven case {i} oracle field {f}")


def test_batch_autotune_restores_state_and_applies_its_choice():
    """ifl_batch_autotune: the trials leave every case's fields as they were, the choice is applied
    to every case, and the batch then continues exactly as single solves with the chosen settings."""
    cs = _uneven(6, 60)
    with Batch(cs) as b:
        b.solve(20)
        before = [_fields(b.cases[i], cs[i]) for i in range(len(cs))]
        r = b.autotune(12)
        assert r["trials"] == 30, r         # 5 tiles x (K 2 | K 4 with Km 0, 4 | K 5 with Km 0, 4, 5)
        assert r["tile"] in (1, 2, 3, 4, 5) and r["sweep_block"] in (2, 4, 5) and r["sweep_block_mom"] in (0, 4, 5)
        for i, c in enumerate(cs):
            assert b.cases[i].cfg.tile == r["tile"] and b.cases[i].cfg.sweep_block == r["sweep_block"]
            for f, v in _fields(b.cases[i], c).items():
                assert_bitwise(v, before[i][f], f"restored case {i} field {f}")
        b.solve(40)
        got = [_fields(b.cases[i], cs[i]) for i in range(len(cs))]
    for i, c in enumerate(cs):
        c2 = c.copy()
        c2.tile, c2.sweep_block, c2.sweep_block_mom = r["tile"], r["sweep_block"], r["sweep_block_mom"]
        with Solver(c2) as s:
            s.solve(20, history=False)
            s.solve(40, history=False)
            want = _fields(s, c2)
        for f in want:
            assert_bitwise(got[i][f], want[f], f"tuned case {i} field {f}")


def test_batch_case_that_diverges_stops_alone():
    """One case of a batch blows up (alpha_p = 1.9): it stops on the device at the iteration its own
    solve reports divergence, the batch returns IFL_ERR_DIVERGED with that case's status, and the
    other cases run on to their own ends, each equal to its single solve (the readback of a chunk
    runs one chunk behind the device, check every 8 iterations)."""
    cs = [cavity(32, 32), cavity(40, 24), channel(48, 16, 30)]
    cs[1].relax[3] = 1.9
    for c in cs:
        c.stop_rule = _abi.STOP_NONE
        c.num_outer_iter = 300
        c.check_every = 8
        c.tile = 3
    with Batch(cs) as b:
        st, reps = b.solve(300)
        got = [_fields(b.cases[i], cs[i]) for i in range(len(cs))]
    assert st == _abi.IFL_ERR_DIVERGED
    assert reps[1]["status"] == _abi.IFL_ERR_DIVERGED and reps[1]["iterations"] < 300
    assert reps[0]["status"] == 0 and reps[0]["iterations"] == 300 and reps[2]["iterations"] == 300
    for i, c in enumerate(cs):
        with Solver(c) as s:
            r = s.solve(300, history=False)
            assert r["iterations"] == reps[i]["iterations"], i
            assert r["status"] == reps[i]["status"], i
            if i != 1:
                for f, v in _fields(s, c).items():
                    assert_bitwise(got[i][f], v, f"case {i} field {f}")
