"""The temperature equation on the GPU (runs on an MI355X): bit-for-bit parity with the FP32
oracle stage by stage and over whole outer iterations (2D tiles and sweeps per launch, 3D,
Jacobi, odd and even sweep counts), batches with energy, state files, and the exact linear
conduction profile."""
import sys

import numpy as np
import pytest

from helpers import STATE_FIELDS, assert_bitwise, cavity, cavity3d, channel, duct3d, load_state, random_state
from incompflow import Batch, IflError, Solver, _abi as A
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FIELDS = STATE_FIELDS + (A.FIELD_TEMP,)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def energy(cfg, n=7, relax=0.8, gam=0.02):
    c = cfg.copy()
    c.energy = 1
    c.n_iter_t = n
    c.relax_t = relax
    c.tol_t = 1e-6
    c.conductivity, c.specific_heat = gam * 4.0, 4.0
    types = [A.TEMP_DIRICHLET, A.TEMP_NEUMANN, A.TEMP_DIRICHLET, A.TEMP_DIRICHLET, A.TEMP_NEUMANN, A.TEMP_DIRICHLET]
    vals = [280.0, 0.0, 330.0, 300.0, 0.0, 310.0]
    for f in range(6):
        c.temp_bc_type[f], c.temp_bc_val[f] = types[f], vals[f]
    c.init[4] = 300.0
    return c


CASES = {
    "channel2d": lambda: energy(channel(40, 12)),
    "cavity2d": lambda: energy(cavity(20, 16), n=8),
    "cavity3d": lambda: energy(cavity3d(), n=5),
    "duct3d": lambda: energy(duct3d(), n=4),
}


def _compare(g, o, cfg, what, extra=()):
    for f in FIELDS + tuple(extra):
        if cfg.dim == 2 and f in (A.FIELD_W, A.FIELD_WF):
            continue
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), f"{what}: {A.FIELD_NAMES[f]}")


@pytest.mark.parametrize("name", sorted(CASES))
def test_energy_stages_bitwise(name):
    cfg = CASES[name]()
    _log(f"[energy stages] {name}")
    g, o = Solver(cfg), Oracle(cfg)
    st = random_state(cfg, seed=5)
    st[A.FIELD_TEMP] = (300.0 + 20.0 * np.random.default_rng(6).uniform(-1, 1, cfg.field_size(A.FIELD_TEMP))).astype(np.float32)
    load_state(g, st)
    load_state(o, st)
    for s in (A.STAGE_INIT_FACES, A.STAGE_MOM_ASSEMBLE):
        g.stage(s)
        o.stage(s)
    _compare(g, o, cfg, "assembly", (A.FIELD_TCOEF, A.FIELD_BT, A.FIELD_MCOEF))
    rg, ro = g.stage(A.STAGE_MOM_RESIDUAL), o.stage(A.STAGE_MOM_RESIDUAL)
    np.testing.assert_allclose(rg[:4], ro[:4], rtol=1e-12, atol=0)
    assert rg[3] > 0
    for s in (A.STAGE_SMOOTH_U, A.STAGE_SMOOTH_V, A.STAGE_SMOOTH_T):
        gg, oo = g.stage(s), o.stage(s)
        assert gg[0] == oo[0]
        assert gg[1] == pytest.approx(oo[1], rel=1e-12, abs=0)
    _compare(g, o, cfg, "after the temperature sweeps")


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("nT", [1, 6, 7])
def test_energy_outer_iterations_bitwise(name, nT):
    cfg = CASES[name]()
    cfg.n_iter_t = nT
    cfg.stop_rule = A.STOP_NONE
    _log(f"[energy outer] {name} n_iter_t={nT}")
    g, o = Solver(cfg), Oracle(cfg)
    rg, ro = g.solve(6), o.solve(6)
    _compare(g, o, cfg, f"{name} n_iter_t={nT}")
    keys = ("res_t", "scaled", "res_u", "res_cont")
    np.testing.assert_allclose([[getattr(r, k) for k in keys] for r in rg["history"]],
                               [[getattr(r, k) for k in keys] for r in ro["history"]], rtol=1e-11, atol=1e-300)
    assert [r.sweeps_t for r in rg["history"]] == [r.sweeps_t for r in ro["history"]]


@pytest.mark.parametrize("tile,block,mblock", [(1, 2, 0), (3, 1, 0), (4, 4, 4), (5, 2, 0), (5, 4, 4), (1, 4, 1),
                                               (5, 5, 5), (2, 5, 0)])
def test_energy_2d_tiles_and_sweep_blocks(tile, block, mblock):
    cfg = energy(cavity(150, 70), n=9)
    cfg.tile, cfg.sweep_block, cfg.sweep_block_mom = tile, block, mblock
    cfg.stop_rule = A.STOP_NONE
    _log(f"[energy tiles] tile={tile} block={block} momentum block={mblock}")
    g, o = Solver(cfg), Oracle(cfg)
    g.solve(4)
    o.solve(4)
    _compare(g, o, cfg, f"tile {tile} block {block}")


def test_energy_3d_two_pcorr_sweeps_per_launch():
    """3D with temperature and two p' sweeps per launch (k_rb_zmarch3d_p2 between the momentum and
    temperature z-marches), several columns and chunks, 5 outer iterations against the oracle."""
    cfg = energy(duct3d(40, 12, 36), n=5)
    cfg.sweep_block = 2
    cfg.n_iter[3] = 9
    cfg.stop_rule = A.STOP_NONE
    g, o = Solver(cfg), Oracle(cfg)
    g.solve(5)
    o.solve(5)
    _compare(g, o, cfg, "3d energy, p' sweep_block 2")


def test_energy_jacobi():
    cfg = CASES["channel2d"]()
    cfg.smoother = A.SMOOTHER_JACOBI
    cfg.stop_rule = A.STOP_NONE
    g, o = Solver(cfg), Oracle(cfg)
    g.solve(4)
    o.solve(4)
    _compare(g, o, cfg, "jacobi")


def test_energy_scaled_stop_same_iteration():
    cfg = energy(cavity(24, 24), n=3, relax=0.3)
    cfg.stop_rule = A.STOP_SCALED
    cfg.tol_residual = 1e-3
    cfg.num_outer_iter = 4000
    g, o = Solver(cfg), Oracle(cfg)
    rg, ro = g.solve(0, history=False), o.solve(4000, history=False)
    assert rg["converged"] and ro["converged"] and rg["iterations"] == ro["iterations"]
    _compare(g, o, cfg, "converged")


def test_energy_batch_equals_single():
    cs = [CASES["channel2d"](), CASES["cavity2d"](), energy(cavity(33, 18, re=20.0))]
    for c in cs:
        c.n_iter_t = 7
        c.tile = 3
        c.stop_rule = A.STOP_NONE
    cs[2].temp_bc_type[A.WEST] = A.TEMP_DIRICHLET
    cs[2].conductivity = 1.0
    with Batch(cs) as b:
        b.solve(12)
        got = [b.cases[i].get(A.FIELD_TEMP, shaped=False) for i in range(len(cs))]
    for i, c in enumerate(cs):
        with Solver(c) as s:
            s.solve(12, history=False)
            assert_bitwise(got[i], s.get(A.FIELD_TEMP, shaped=False), f"batch case {i}")
    bad = [cs[0], cs[1].copy()]
    bad[1].n_iter_t = 3
    with pytest.raises(IflError, match="n_iter_t"):
        Batch(bad)


def test_energy_state_resume(tmp_path):
    cfg = CASES["duct3d"]()
    cfg.stop_rule = A.STOP_SCALED
    cfg.tol_residual = 1e-9
    a = Solver(cfg)
    a.solve(3, history=False)
    path = str(tmp_path / "e.bin")
    a.save_state(path)
    ra = a.solve(6)
    b = Solver(cfg)
    b.load_state(path)
    rb = b.solve(6)
    assert [(r.res_t, r.scaled) for r in ra["history"]] == [(r.res_t, r.scaled) for r in rb["history"]]
    assert_bitwise(a.get(A.FIELD_TEMP), b.get(A.FIELD_TEMP), "temp")


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_conduction_profile_on_device(dim):
    """Fluid at rest between Dirichlet walls (south 373, north 293), other faces adiabatic: the
    converged temperature is the exact linear profile up to FP32 stagnation (a Gauss-Seidel
    update below half an ulp of T, 1.5e-5 K at 300 K, rounds away; the local residuals left then
    add up over the cells between the walls to ~5e-5 of T) and equals the FP32 oracle bit for bit."""
    base = cavity(32, 48) if dim == 2 else cavity3d(16, 24, 8)
    base.bc_val[A.NORTH][0] = 0.0
    base.bc_val[A.TOP][0] = 0.0
    c = energy(base, n=40000, relax=1.0, gam=0.5)
    c.tol_t = 0.0
    for f in range(6):
        c.temp_bc_type[f], c.temp_bc_val[f] = A.TEMP_NEUMANN, 0.0
    c.temp_bc_type[A.NORTH], c.temp_bc_val[A.NORTH] = A.TEMP_DIRICHLET, 293.0
    c.temp_bc_type[A.SOUTH], c.temp_bc_val[A.SOUTH] = A.TEMP_DIRICHLET, 373.0
    g = Solver(c)
    g.stage(A.STAGE_MOM_ASSEMBLE)
    g.stage(A.STAGE_SMOOTH_T)
    T = g.get(A.FIELD_TEMP).astype(np.float64)
    H = c.ymax - c.ymin
    y = (np.arange(c.ny) + 0.5) * H / c.ny
    exact = 373.0 + (293.0 - 373.0) * y / H
    np.testing.assert_allclose(T, np.broadcast_to(exact[None, :, None], T.shape), rtol=1e-4)
    o = Oracle(c)
    o.stage(A.STAGE_MOM_ASSEMBLE)
    o.stage(A.STAGE_SMOOTH_T)
    assert_bitwise(g.get(A.FIELD_TEMP), o.get(A.FIELD_TEMP), "conduction")


def test_energy_off_has_no_temperature_matrix():
    cfg = cavity(16, 12)
    g = Solver(cfg)
    g.solve(3)
    assert np.all(g.get(A.FIELD_TEMP) == np.float32(cfg.init[4]))
    with pytest.raises(IflError, match="energy"):
        g.get(A.FIELD_TCOEF)
    with pytest.raises(IflError, match="energy"):
        g.stage(A.STAGE_SMOOTH_T)
