"""The multigrid p' solve on 3D grids (IFL_PSOLVE_MG, mg3d.hip) against the FP32 oracle, bit for bit.

Grids cover a single coarsening (sides below 8), several levels with odd sides (the last aggregate of
an odd side holds one cell), a duct with an inlet and an outlet (the outlet ghost coupling of level 0
and its sums on the coarse levels) and the smallest grid the library accepts.
"""
import sys

import pytest

from helpers import assert_bitwise, cavity3d, duct3d, load_state, random_state
from incompflow import Solver, _abi
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

A = _abi
FIELDS = (A.FIELD_U, A.FIELD_V, A.FIELD_W, A.FIELD_P, A.FIELD_UF, A.FIELD_VF, A.FIELD_WF)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mg(cfg, cycles=1, nu=2, ncoarse=7, over=1.4):
    cfg.p_solver = A.PSOLVE_MG
    cfg.p_refine = 0
    cfg.mg_cycles = cycles
    cfg.mg_sweeps = nu
    cfg.mg_coarse_sweeps = ncoarse
    cfg.mg_over = over
    return cfg


def _outlet_duct(nx, ny, nz):
    c = duct3d(nx, ny, nz)
    c.bc_type[A.NORTH] = A.OUTLET
    return c


GRIDS = {
    "cavity6x5x7": lambda: cavity3d(6, 5, 7),           # one coarse level of 3x3x4 cells
    "cavity4x5x6": lambda: cavity3d(4, 5, 6),           # the smallest grid: one coarse level of 2x3x3
    "cavity37x29x23": lambda: cavity3d(37, 29, 23),     # four levels, odd sides
    "duct41x18x13": lambda: _outlet_duct(41, 18, 13),   # inlet, outlets on east and north
    "cavity64": lambda: cavity3d(64, 64, 64),           # five levels
}


@pytest.mark.parametrize("name", list(GRIDS))
@pytest.mark.parametrize("nu,cycles", [(2, 1), (1, 2), (3, 1)])
def test_mg3_p_solve_stage_bitwise(name, nu, cycles):
    cfg = _mg(GRIDS[name](), cycles=cycles, nu=nu)
    _log(f"[mg3 stage] {name} nu={nu} cycles={cycles}")
    g, o = Solver(cfg), Oracle(cfg)
    st = random_state(cfg, seed=13)
    load_state(g, st)
    load_state(o, st)
    for sid in (A.STAGE_MOM_ASSEMBLE, A.STAGE_SMOOTH_U, A.STAGE_SMOOTH_V, A.STAGE_SMOOTH_W, A.STAGE_RHIE_CHOW,
                A.STAGE_PCORR_ASSEMBLE):
        o.stage(sid)
        g.stage(sid)
    assert_bitwise(g.get(A.FIELD_PCOEF, shaped=False), o.get(A.FIELD_PCOEF, shaped=False), "p' matrix")
    og = o.stage(A.STAGE_SMOOTH_P)
    gg = g.stage(A.STAGE_SMOOTH_P)
    assert gg[0] == og[0] == cycles * 2 * nu
    assert_bitwise(g.get(A.FIELD_PCORR, shaped=False), o.get(A.FIELD_PCORR, shaped=False), f"{name}: p'")
    o.stage(A.STAGE_UPDATE)
    g.stage(A.STAGE_UPDATE)
    for f in FIELDS:
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), f"{name}: {A.FIELD_NAMES[f]}")
    g.close()


@pytest.mark.parametrize("nu,cycles", [(2, 1), (3, 2)])
def test_mg3_zmarch_levels_bitwise(nu, cycles):
    """A grid large enough for the z-march smoother on level 0 (at least 512 columns of 32x8x32):
    odd sides and ragged columns in x, y and z; pairs of sweeps march, an odd last sweep and the
    prolongation it then needs run level-wide."""
    cfg = _mg(_outlet_duct(161, 163, 167), cycles=cycles, nu=nu)
    _log(f"[mg3 z-march] duct 161x163x167 nu={nu} cycles={cycles}")
    g, o = Solver(cfg), Oracle(cfg)
    st = random_state(cfg, seed=17)
    load_state(g, st)
    load_state(o, st)
    for sid in (A.STAGE_MOM_ASSEMBLE, A.STAGE_RHIE_CHOW, A.STAGE_PCORR_ASSEMBLE):
        o.stage(sid)
        g.stage(sid)
    og = o.stage(A.STAGE_SMOOTH_P)
    gg = g.stage(A.STAGE_SMOOTH_P)
    assert gg[0] == og[0] == cycles * 2 * nu
    assert_bitwise(g.get(A.FIELD_PCORR, shaped=False), o.get(A.FIELD_PCORR, shaped=False), "p'")
    g.close()


@pytest.mark.parametrize("name,over", [("cavity37x29x23", 1.4), ("duct41x18x13", 1.0)])
def test_mg3_outer_iterations_bitwise(name, over):
    """20 outer iterations from rest through the captured graph, bit for bit, with the same records."""
    cfg = _mg(GRIDS[name](), over=over)
    cfg.stop_rule = A.STOP_NONE
    _log(f"[mg3 outer] {name} over={over}")
    g, o = Solver(cfg), Oracle(cfg)
    rg = g.solve(20)
    ro = o.solve(20)
    assert rg["status"] == ro["status"] == 0
    assert rg["iterations"] == ro["iterations"] == 20
    for f in FIELDS:
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), f"{name}: {A.FIELD_NAMES[f]}")
    for a, b in zip(rg["history"], ro["history"]):
        assert a.sweeps[3] == b.sweeps[3] == 4
        assert a.scaled == pytest.approx(b.scaled, rel=1e-11)
    g.close()


def test_mg3_cavity_converges_like_the_oracle_and_faster_than_sweeps():
    """The 3D cavity at 32^3 (Re 400) to a 1e-4 scaled residual: the device stops at the oracle's
    iteration with the oracle's fields, in fewer outer iterations than the reference's p' sweeps."""
    base = A.cavity3d_config(n=32, tol_residual=1e-4, max_iter=3000)
    base.p_refine = 0
    cfg = _mg(base.copy(), ncoarse=4)
    _log("[mg3 converge] cavity 32^3")
    g, o = Solver(cfg), Oracle(cfg)
    rg = g.solve(3000, history=False)
    ro = o.solve(3000, history=False)
    assert rg["converged"] and ro["converged"]
    assert rg["iterations"] == ro["iterations"]
    for f in (A.FIELD_U, A.FIELD_V, A.FIELD_W, A.FIELD_P):
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), A.FIELD_NAMES[f])
    g.close()
    gs = Solver(base)
    rs = gs.solve(3000, history=False)
    gs.close()
    assert rs["converged"]
    assert rg["iterations"] < rs["iterations"], (rg["iterations"], rs["iterations"])


def test_mg3_timing_hook_and_autotune_leave_the_state():
    """ifl_time_kernel(IFL_KERNEL_PSOLVE) and the p' solver autotuner restore the arena: the solve
    afterwards equals an untouched one bit for bit."""
    cfg = _mg(cavity3d(24, 20, 18))
    a, b = Solver(cfg), Solver(cfg)
    a.solve(5)
    b.solve(5)
    a.time_kernel(A.KERNEL_PSOLVE, reps=3)
    a.solve(20)                 # 25 iterations: the tuner advances younger cases to 20 first
    b.solve(20)
    for f in FIELDS:
        assert_bitwise(a.get(f, shaped=False), b.get(f, shaped=False), A.FIELD_NAMES[f])
    before = {f: a.get(f, shaped=False) for f in FIELDS}
    r = a.autotune(A.TUNE_PSOLVE, trial_iters=40)
    assert r["trials"] == 5 and r["p_solver"] in (A.PSOLVE_SWEEPS, A.PSOLVE_MG)
    for f in FIELDS:
        assert_bitwise(a.get(f, shaped=False), before[f], f"restored {A.FIELD_NAMES[f]}")
    a.close()
    b.close()
