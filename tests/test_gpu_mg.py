"""The multigrid p' solve on the device (IFL_PSOLVE_MG) against the FP32 oracle, bit for bit.

Grid sizes cover the single-workgroup bottom kernel alone (small grids), the tiled down/up kernels
above it with ragged last tiles and odd sides, and two coarsening passes (more than six levels);
cases cover walls only (a singular p' system) and inlets/outlets (the outlet ghost coupling).
Also: the timing hook leaves the state as it found it, the autotuner's p' solver choice, and the
fail_if_not_converged status.
"""
import sys

import numpy as np
import pytest

from helpers import assert_bitwise, cavity, channel, load_state, random_state
from incompflow import Solver, _abi
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

A = _abi


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mg(cfg, cycles=1, nu=2, ncoarse=20, over=1.4):
    cfg.p_solver = A.PSOLVE_MG
    cfg.p_refine = 0
    cfg.mg_cycles = cycles
    cfg.mg_sweeps = nu
    cfg.mg_coarse_sweeps = ncoarse
    cfg.mg_over = over
    return cfg


def _outlet_box(nx, ny):
    c = cavity(nx, ny)
    c.bc_type[A.WEST] = A.INLET
    c.bc_val[A.WEST][0] = 0.5
    c.bc_type[A.EAST] = A.OUTLET
    c.bc_type[A.NORTH] = A.OUTLET
    return c


GRIDS = {
    "cavity48x40": lambda: cavity(48, 40),          # bottom kernel only
    "cavity200x130": lambda: cavity(200, 130),      # down/up tiles + bottom, ragged tiles
    "box257x67": lambda: _outlet_box(257, 67),      # odd sides, outlet terms
    "channel512x128": lambda: channel(512, 128),    # the reference's channel, inlet + outlet
    "cavity1100x900": lambda: cavity(1100, 900),    # two coarsening passes (10 levels)
}


@pytest.mark.parametrize("name", list(GRIDS))
@pytest.mark.parametrize("nu,cycles", [(2, 1), (1, 2), (3, 1)])
def test_mg_p_solve_stage_bitwise(name, nu, cycles):
    cfg = _mg(GRIDS[name](), cycles=cycles, nu=nu, ncoarse=7)
    _log(f"[mg stage] {name} nu={nu} cycles={cycles}")
    g, o = Solver(cfg), Oracle(cfg)
    st = random_state(cfg, seed=11)
    load_state(g, st)
    load_state(o, st)
    for sid in (A.STAGE_MOM_ASSEMBLE, A.STAGE_SMOOTH_U, A.STAGE_SMOOTH_V, A.STAGE_RHIE_CHOW, A.STAGE_PCORR_ASSEMBLE):
        o.stage(sid)
        g.stage(sid)
    assert_bitwise(g.get(A.FIELD_PCOEF, shaped=False), o.get(A.FIELD_PCOEF, shaped=False), "p' matrix")
    og = o.stage(A.STAGE_SMOOTH_P)
    gg = g.stage(A.STAGE_SMOOTH_P)
    assert gg[0] == og[0] == cycles * 2 * nu
    assert_bitwise(g.get(A.FIELD_PCORR, shaped=False), o.get(A.FIELD_PCORR, shaped=False), f"{name}: p'")
    o.stage(A.STAGE_UPDATE)
    g.stage(A.STAGE_UPDATE)
    for f in (A.FIELD_U, A.FIELD_V, A.FIELD_P, A.FIELD_UF, A.FIELD_VF):
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), f"{name}: {A.FIELD_NAMES[f]}")
    g.close()


@pytest.mark.parametrize("name,over", [("cavity200x130", 1.4), ("box257x67", 1.0), ("box257x67", 1.4)])
def test_mg_outer_iterations_bitwise(name, over):
    """25 outer iterations from rest, bit for bit.  On the outlet box a coarse-correction factor of
    1.4 is too strong: device and oracle diverge together, at the same iteration, with the same
    records up to it (a p' solve of controlled strength is a tuning choice)."""
    cfg = _mg(GRIDS[name](), over=over)
    cfg.stop_rule = A.STOP_NONE
    _log(f"[mg outer] {name}")
    g, o = Solver(cfg), Oracle(cfg)
    rg = g.solve(25)
    ro = o.solve(25)
    assert rg["iterations"] == ro["iterations"]
    assert rg["status"] == ro["status"]
    if ro["status"] == 0:
        assert rg["iterations"] == 25
        for f in (A.FIELD_U, A.FIELD_V, A.FIELD_P, A.FIELD_UF, A.FIELD_VF):
            assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), f"{name}: {A.FIELD_NAMES[f]}")
    else:
        assert (name, over) == ("box257x67", 1.4) and ro["status"] == A.IFL_ERR_DIVERGED
    for a, b in zip(rg["history"][:-1], ro["history"][:-1]):
        assert a.sweeps[3] == b.sweeps[3] == 4
        assert a.scaled == pytest.approx(b.scaled, rel=1e-11)
    g.close()


def test_mg_cavity_converges_like_the_oracle_and_faster_than_sweeps():
    """128x128 cavity to a 1e-4 scaled residual: the device stops at the oracle's iteration with
    the oracle's fields, in well under half the outer iterations of the reference's p' sweeps."""
    base = A.cavity_config(n=128, p_refine=0, tol_residual=1e-4, max_iter=6000)
    cfg = _mg(base.copy(), over=1.8)
    _log("[mg converge] cavity 128")
    g, o = Solver(cfg), Oracle(cfg)
    rg = g.solve(6000, history=False)
    ro = o.solve(6000, history=False)
    assert rg["converged"] and ro["converged"]
    assert rg["iterations"] == ro["iterations"]
    for f in (A.FIELD_U, A.FIELD_V, A.FIELD_P):
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), A.FIELD_NAMES[f])
    g.close()
    gs = Solver(base)
    rs = gs.solve(6000, history=False)
    gs.close()
    assert rs["converged"]
    assert rg["iterations"] < 0.5 * rs["iterations"], (rg["iterations"], rs["iterations"])


@pytest.mark.parametrize("kernel", [A.KERNEL_MOM_SWEEP, A.KERNEL_UPDATE, A.KERNEL_SWEEP, A.KERNEL_PSOLVE])
@pytest.mark.parametrize("p_solver", [A.PSOLVE_SWEEPS, A.PSOLVE_MG])
def test_timing_a_kernel_leaves_the_solve_unchanged(kernel, p_solver):
    """ifl_time_kernel restores the arena: a solve after timing (odd momentum launch counts,
    sweep_block_mom 4 on 10 sweeps) equals a fresh solve bit for bit."""
    cfg = cavity(96, 80)
    cfg.sweep_block = 4
    cfg.sweep_block_mom = 4
    if p_solver == A.PSOLVE_MG:
        _mg(cfg)
    _log(f"[time kernel] kernel={kernel} p_solver={p_solver}")
    a, b = Solver(cfg), Solver(cfg)
    a.solve(7)
    b.solve(7)
    a.time_kernel(kernel, reps=5)
    a.solve(9)
    b.solve(9)
    for f in (A.FIELD_U, A.FIELD_V, A.FIELD_P, A.FIELD_UF, A.FIELD_VF):
        assert_bitwise(a.get(f, shaped=False), b.get(f, shaped=False), A.FIELD_NAMES[f])
    a.close()
    b.close()


def test_autotune_chooses_a_p_solver_and_the_solve_uses_it():
    """IFL_TUNE_PSOLVE: five trials (the sweeps, three coarse-correction factors, the final timing),
    the state restored bit for bit, and the solve then continues with the choice exactly as the
    oracle continues from the same fields with the same settings."""
    cfg = A.cavity_config(n=128, p_refine=0, tol_residual=1e-4, max_iter=4000)
    _log("[autotune psolve] cavity 128")
    g = Solver(cfg)
    g.solve(40, history=False)
    fields = (A.FIELD_U, A.FIELD_V, A.FIELD_P, A.FIELD_UF, A.FIELD_VF)
    before = {f: g.get(f, shaped=False) for f in fields}
    r = g.autotune(A.TUNE_PSOLVE, trial_iters=150)
    assert r["trials"] == 5
    assert r["p_solver"] in (A.PSOLVE_SWEEPS, A.PSOLVE_MG)
    assert g.cfg.p_solver == r["p_solver"]
    if r["p_solver"] == A.PSOLVE_MG:
        assert r["mg_over"] in (1.0, 1.4, 1.8) and r["mg_cycles"] == 1
    for f in fields:
        assert_bitwise(g.get(f, shaped=False), before[f], f"restored {A.FIELD_NAMES[f]}")
    c2 = g.cfg.copy()
    c2.stop_rule = A.STOP_NONE          # the oracle has no normalisers (its counter starts at 40)
    o = Oracle(c2)
    for f in fields:
        o.set(f, before[f])
    o.set_iter(40)
    rg = g.solve(5, history=False)
    ro = o.solve(5, history=False)
    assert rg["iterations"] == ro["iterations"] == 5
    for f in fields:
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), f"continued {A.FIELD_NAMES[f]}")
    g.close()


def test_fail_if_not_converged_status():
    cfg = A.cavity_config(n=32, p_refine=0, tol_residual=1e-12, max_iter=30)
    cfg.fail_if_not_converged = 1
    g = Solver(cfg)
    r = g.solve(30, history=False)
    assert r["status"] == A.IFL_ERR_NOT_CONVERGED and not r["converged"] and r["iterations"] == 30
    g.close()
    cfg.tol_residual = 0.5
    g = Solver(cfg)
    r = g.solve(30, history=False)
    assert r["status"] == 0 and r["converged"]
    g.close()


@pytest.mark.parametrize("case", ["cavity128", "channel256x64_profile", "cavity3d_40"])
def test_autotune_by_time_to_target_leaves_the_state(case):
    """IFL_TUNE_PSOLVE_TARGET: four proxy solves (the sweeps, three coarse-correction factors) on a
    coarsened copy of the case, inlet profile averaged onto it; the handle's fields and iteration
    count are untouched, the choice is applied, and the solve then continues with it."""
    if case == "cavity128":
        cfg = A.cavity_config(n=128, p_refine=1, tol_residual=1e-4, max_iter=6000)
        prof = None
    elif case == "channel256x64_profile":
        cfg = A.adjoint_channel_config(256, 64)
        cfg.tol_residual = 1e-4
        prof = np.array(A.adjoint_inlet_profile(cfg))
    else:
        cfg = A.cavity3d_config(n=40, tol_residual=1e-4, max_iter=3000)
        cfg.p_refine = 0
        prof = None
    _log(f"[autotune target] {case}")
    g = Solver(cfg)
    if prof is not None:
        g.set_face_profile(A.WEST, prof)
    g.solve(10, history=False)
    fields = (A.FIELD_U, A.FIELD_V, A.FIELD_P, A.FIELD_UF, A.FIELD_VF)
    before = {f: g.get(f, shaped=False) for f in fields}
    r = g.autotune(A.TUNE_PSOLVE_TARGET)
    assert r["trials"] == 4
    assert r["p_solver"] in (A.PSOLVE_SWEEPS, A.PSOLVE_MG) and g.cfg.p_solver == r["p_solver"]
    for f in fields:
        assert_bitwise(g.get(f, shaped=False), before[f], f"untouched {A.FIELD_NAMES[f]}")
    rr = g.solve(5, history=True)
    assert rr["iterations"] == 5 and rr["status"] == 0
    if r["p_solver"] == A.PSOLVE_MG:
        assert all(x.sweeps[3] == 4 for x in rr["history"])
    g.close()
