"""Parity of the gfx950 library with the FP32 CPU oracle (runs on an MI355X).

Stage outputs must be bit-identical to the oracle's: both sides evaluate kernels.cuh's
expressions with the same float/double promotions and no fused multiply-add, and red-black
ordering makes each half sweep order independent.  FP64 norms are compared at 1e-12 relative
(their summation order differs).
"""
import os
import sys

import numpy as np
import pytest

from helpers import CASES, STATE_FIELDS, assert_bitwise, cavity, load_state, max_rel, random_state
from incompflow import _abi
from incompflow import Solver, IflError
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

OUT_FIELDS = STATE_FIELDS + (_abi.FIELD_MCOEF, _abi.FIELD_PCOEF, _abi.FIELD_BU, _abi.FIELD_BV, _abi.FIELD_BP)

STAGES = [
    ("FACE_BC", _abi.STAGE_FACE_BC),
    ("MOM_ASSEMBLE", _abi.STAGE_MOM_ASSEMBLE),
    ("MOM_RESIDUAL", _abi.STAGE_MOM_RESIDUAL),
    ("SMOOTH_U", _abi.STAGE_SMOOTH_U),
    ("SMOOTH_V", _abi.STAGE_SMOOTH_V),
    ("SMOOTH_W", _abi.STAGE_SMOOTH_W),
    ("RHIE_CHOW", _abi.STAGE_RHIE_CHOW),
    ("PCORR_ASSEMBLE", _abi.STAGE_PCORR_ASSEMBLE),
    ("SMOOTH_P", _abi.STAGE_SMOOTH_P),
    ("UPDATE", _abi.STAGE_UPDATE),
]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare_all(g, o, cfg, what):
    for f in OUT_FIELDS:
        if cfg.dim == 2 and f in (_abi.FIELD_W, _abi.FIELD_WF):
            continue
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), f"{what}: field {_abi.FIELD_NAMES[f]}")


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("p_refine", [0, 1])
def test_stage_by_stage_bitwise(name, p_refine):
    cfg = CASES[name]()
    cfg.p_refine = p_refine
    _log(f"[stage parity] case={name} p_refine={p_refine} grid={cfg.nx}x{cfg.ny}x{cfg.nz}")
    g = Solver(cfg)
    o = Oracle(cfg)
    st = random_state(cfg, seed=7)
    load_state(g, st)
    load_state(o, st)
    g.stage(_abi.STAGE_INIT_FACES)
    o.stage(_abi.STAGE_INIT_FACES)
    _compare_all(g, o, cfg, "INIT_FACES")
    for label, sid in STAGES:
        if sid == _abi.STAGE_SMOOTH_W and cfg.dim == 2:
            continue
        og = o.stage(sid)
        gg = g.stage(sid)
        _compare_all(g, o, cfg, label)
        if sid in (_abi.STAGE_SMOOTH_U, _abi.STAGE_SMOOTH_V, _abi.STAGE_SMOOTH_W, _abi.STAGE_SMOOTH_P):
            assert gg[0] == og[0], f"{label}: sweeps {gg[0]} vs {og[0]}"
            assert gg[1] == pytest.approx(og[1], rel=1e-12, abs=0)
        elif sid in (_abi.STAGE_MOM_RESIDUAL, _abi.STAGE_PCORR_ASSEMBLE, _abi.STAGE_UPDATE):
            np.testing.assert_allclose(gg, og, rtol=1e-12, atol=0, err_msg=label)


def _rec_array(history, keys=("p_norm", "u_norm", "v_norm", "w_norm", "uf_norm", "vf_norm", "wf_norm",
                              "mass_norm", "res_cont", "res_u", "res_v", "res_w", "scaled")):
    return np.array([[getattr(r, k) for k in keys] for r in history])


@pytest.mark.parametrize("name", sorted(CASES))
def test_outer_iterations_graph_bitwise(name):
    """Five outer iterations through the captured hipGraph equal the oracle's loop."""
    cfg = CASES[name]()
    cfg.stop_rule = _abi.STOP_NONE
    _log(f"[outer parity] case={name}")
    g = Solver(cfg)
    o = Oracle(cfg)
    rg = g.solve(5)
    ro = o.solve(5)
    assert rg["iterations"] == ro["iterations"] == 5
    _compare_all(g, o, cfg, f"{name} after 5 iterations")
    np.testing.assert_allclose(_rec_array(rg["history"]), _rec_array(ro["history"]), rtol=1e-11, atol=1e-300)
    assert [tuple(r.sweeps) for r in rg["history"]] == [tuple(r.sweeps) for r in ro["history"]]


def test_graph_equals_stage_path():
    cfg = CASES["cavity2d"]()
    a = Solver(cfg)
    b = Solver(cfg)
    a.solve(4)
    for _ in range(4):
        b.stage(_abi.STAGE_OUTER)
    for f in (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF):
        assert_bitwise(a.get(f), b.get(f), _abi.FIELD_NAMES[f])


def test_determinism_two_runs():
    cfg = CASES["channel2d"]()
    runs = []
    for _ in range(2):
        g = Solver(cfg)
        g.solve(20)
        runs.append([g.get(f) for f in (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P)])
    for a, b in zip(*runs):
        assert_bitwise(a, b, "repeat run")


def test_jacobi_smoother_parity():
    cfg = CASES["cavity2d"]()
    cfg.smoother = _abi.SMOOTHER_JACOBI
    g = Solver(cfg)
    o = Oracle(cfg)
    g.solve(4)
    o.solve(4)
    _compare_all(g, o, cfg, "jacobi")


def test_inlet_profile_parity():
    cfg = CASES["channel2d"]()
    prof = 0.001 * (1.0 + 0.3 * np.sin(np.arange(cfg.ny)))
    g = Solver(cfg)
    o = Oracle(cfg)
    g.set_face_profile(_abi.WEST, prof)
    o.set_face_profile(_abi.WEST, prof)
    g.solve(6)
    o.solve(6)
    _compare_all(g, o, cfg, "profile")


def test_converged_cavity_same_iterations():
    """Scaled-residual stop: same iteration count and bit-identical converged fields."""
    cfg = _abi.cavity_config(n=32, p_refine=0, tol_residual=1e-5, max_iter=5000)
    _log("[converged] cavity 32x32 to 1e-5")
    g = Solver(cfg)
    o = Oracle(cfg)
    rg = g.solve(5000, history=False)
    ro = o.solve(5000, history=False)
    assert rg["converged"] and ro["converged"]
    assert rg["iterations"] == ro["iterations"]
    _compare_all(g, o, cfg, "converged cavity")


def test_channel_c1_3000_iterations():
    """Plain method, 512x128 channel, 3000 outer iterations, field by field vs the FP32 oracle."""
    cfg = _abi.channel_config(512, 128, 3000)
    _log("[C1] channel 512x128, 3000 iterations")
    g = Solver(cfg)
    rg = g.solve(3000, history=False)
    assert rg["iterations"] == 3000
    o = Oracle(cfg)
    ro = o.solve(3000, history=False)
    assert ro["iterations"] == 3000
    for f in (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P):
        a, b = g.get(f), o.get(f)
        assert max_rel(a, b) <= 1e-4, _abi.FIELD_NAMES[f]
        assert_bitwise(a, b, f"C1 {_abi.FIELD_NAMES[f]}")


def test_nan_initial_field_reports_divergence():
    cfg = CASES["cavity2d"]()
    g = Solver(cfg)
    u = g.get(_abi.FIELD_U)
    u[0, 3, 4] = np.nan
    g.set(_abi.FIELD_U, u)
    r = g.solve(10)
    assert r["status"] == _abi.IFL_ERR_DIVERGED
    assert r["iterations"] == 1


@pytest.mark.parametrize("which", ["u", "p"])
@pytest.mark.parametrize("name", ["cavity2d", "channel2d", "cavity3d"])
def test_over_relaxation_diverges_at_the_oracles_iteration(name, which):
    """alpha = 1.9 blows the iteration up; the device must return IFL_ERR_DIVERGED at the same
    outer iteration as the FP32 oracle (their iterates are the same bits until then)."""
    cfg = CASES[name]()
    cfg.stop_rule = _abi.STOP_NONE
    if which == "u":
        for q in range(3):
            cfg.relax[q] = 1.9
    else:
        cfg.relax[3] = 1.9
    _log(f"[divergence] case={name} alpha_{which}=1.9 grid={cfg.nx}x{cfg.ny}x{cfg.nz}")
    g = Solver(cfg)
    r = g.solve(3000)
    ro = Oracle(cfg).solve(3000, history=False)
    assert ro["status"] == _abi.IFL_ERR_DIVERGED
    assert r["status"] == _abi.IFL_ERR_DIVERGED
    assert r["iterations"] == ro["iterations"] < 200


def test_vtk_output_format(tmp_path):
    cfg = CASES["cavity2d"]()
    g = Solver(cfg)
    g.solve(3)
    path = tmp_path / "out" / "case.vtk"
    g.write_vtk(str(path))
    text = path.read_text()
    lines = text.split("\n")
    assert lines[0] == "# vtk DataFile Version 3.0"
    assert lines[3] == "DATASET RECTILINEAR_GRID"
    assert lines[4] == f"DIMENSIONS {cfg.nx + 1} {cfg.ny + 1} {cfg.nz + 1}"
    assert lines[5] == f"X_COORDINATES {cfg.nx + 1} float"
    xs = np.array(lines[6].split(), dtype=np.float64)
    np.testing.assert_allclose(xs, np.linspace(cfg.xmin, cfg.xmax, cfg.nx + 1), atol=1e-6)
    ncell = cfg.nx * cfg.ny * cfg.nz
    assert f"CELL_DATA {ncell}" in lines
    i = lines.index(f"velocity 3 {ncell} float")
    vel = np.array(lines[i + 1].split(), dtype=np.float64).reshape(-1, 3)
    u = g.get(_abi.FIELD_U).ravel()
    np.testing.assert_allclose(vel[:, 0], u, rtol=1e-5, atol=1e-6)
    assert not text.endswith("\n")   # Solver.cu:361-366 writes no final newline


def test_bad_config_rejected():
    cfg = CASES["cavity2d"]()
    cfg.nx = 2
    with pytest.raises(IflError):
        Solver(cfg)


@pytest.mark.parametrize("tile", [_abi.TILE_64x16, _abi.TILE_64x8, _abi.TILE_32x8, _abi.TILE_32x4, _abi.TILE_128x32])
@pytest.mark.parametrize("shape", [(70, 37), (33, 18), (150, 45)])
def test_fused_sweep_tiles_ragged_grids(tile, shape):
    """Every tile variant of the fused sweep on grids that are not tile multiples."""
    cfg = _abi.channel_config(shape[0], shape[1], 6)
    cfg.tile = tile
    _log(f"[tiles] tile={tile} grid={shape}")
    g = Solver(cfg)
    o = Oracle(cfg)
    g.solve(6)
    o.solve(6)
    _compare_all(g, o, cfg, f"tile {tile} {shape}")


def test_odd_and_unequal_sweep_counts():
    """nIter_u != nIter_v and odd counts exercise the partially fused launches and copy-backs."""
    cfg = CASES["cavity2d"]()
    cfg.n_iter[0], cfg.n_iter[1], cfg.n_iter[3] = 3, 6, 7
    cfg.p_refine = 1
    g = Solver(cfg)
    o = Oracle(cfg)
    rg = g.solve(5)
    ro = o.solve(5)
    _compare_all(g, o, cfg, "odd sweeps")
    assert [tuple(r.sweeps) for r in rg["history"]] == [tuple(r.sweeps) for r in ro["history"]]


@pytest.mark.parametrize("block", [2, 4, 5])
@pytest.mark.parametrize("tile", [_abi.TILE_64x16, _abi.TILE_32x4, _abi.TILE_128x32])
@pytest.mark.parametrize("name", ["channel2d", "cavity2d"])
def test_two_sweeps_per_launch(tile, name, block):
    """sweep_block=2, 4 and 5 (two, four or five red-black sweeps per launch over a 4-, 8- or
    10-cell halo; the momentum equations keep 2) vs the oracle's blocked stop rule, with odd and
    unequal sweep counts and the p' refinement."""
    cfg = CASES[name]()
    cfg.sweep_block = block
    cfg.tile = tile
    cfg.n_iter[0], cfg.n_iter[1], cfg.n_iter[3] = 5, 4, 7
    cfg.p_refine = 1
    _log(f"[block{block}] case={name} tile={tile}")
    g = Solver(cfg)
    o = Oracle(cfg)
    rg = g.solve(6)
    ro = o.solve(6)
    _compare_all(g, o, cfg, f"block2 {name}")
    assert [tuple(r.sweeps) for r in rg["history"]] == [tuple(r.sweeps) for r in ro["history"]]
    np.testing.assert_allclose(_rec_array(rg["history"]), _rec_array(ro["history"]), rtol=1e-11, atol=1e-300)


@pytest.mark.parametrize("tile,shape", [(_abi.TILE_64x16, (150, 70)), (_abi.TILE_32x8, (40, 30)),
                                        (_abi.TILE_128x32, (400, 128))])
@pytest.mark.parametrize("mom_block", [1, 4, 5])
def test_momentum_sweep_block(tile, shape, mom_block):
    """sweep_block_mom (momentum sweeps per launch) 1, 4 or 5 with four p' sweeps per launch: the
    blocked stop rule of the oracle, odd and unequal sweep counts, interior and boundary tiles."""
    cfg = cavity(*shape)
    cfg.sweep_block, cfg.sweep_block_mom, cfg.tile = 4, mom_block, tile
    cfg.n_iter[0], cfg.n_iter[1], cfg.n_iter[3] = 7, 10, 12
    cfg.tol[0] = 1e-2
    _log(f"[mom block {mom_block}] tile={tile} shape={shape}")
    g, o = Solver(cfg), Oracle(cfg)
    rg, ro = g.solve(5), o.solve(5)
    _compare_all(g, o, cfg, f"momentum block {mom_block}")
    assert [tuple(r.sweeps) for r in rg["history"]] == [tuple(r.sweeps) for r in ro["history"]]


@pytest.mark.parametrize("tile,shape,block", [(_abi.TILE_128x32, (4224, 1024), 5), (_abi.TILE_64x16, (2112, 1024), 4)])
def test_many_tiles_stop_decisions(tile, shape, block):
    """More tiles than one pass of prefetched partial records: 1056 tiles of 128x32 (1024-thread
    workgroups) or 2112 of 64x16 (256 threads), so every wave of a sweep's stop decision holds
    records and, with 64x16, the loop over the records past the prefetched ones runs; loose sweep
    tolerances make stop tests trigger inside blocks (the v sweeps stop early)."""
    cfg = cavity(*shape)
    cfg.tile, cfg.sweep_block, cfg.sweep_block_mom = tile, block, block
    cfg.n_iter[0], cfg.n_iter[1], cfg.n_iter[3] = 7, 10, 12
    cfg.tol[0], cfg.tol[1], cfg.tol[3] = 2e-2, 2e-2, 0.25
    _log(f"[many tiles] tile={tile} grid={shape} block={block}")
    g, o = Solver(cfg), Oracle(cfg)
    rg, ro = g.solve(3), o.solve(3)
    _compare_all(g, o, cfg, f"many tiles {tile}")
    sweeps = [tuple(r.sweeps) for r in rg["history"]]
    assert sweeps == [tuple(r.sweeps) for r in ro["history"]]
    assert any(sw[0] < cfg.n_iter[0] or sw[1] < cfg.n_iter[1] or sw[3] < cfg.n_iter[3] for sw in sweeps), sweeps


@pytest.mark.parametrize("block", [2, 4, 5])
@pytest.mark.parametrize("tile,shape", [(_abi.TILE_64x16, (256, 96)), (_abi.TILE_128x32, (400, 128))])
def test_sweep_blocks_on_interior_tiles_with_refinement(block, tile, shape):
    """A grid with interior tiles (their ring lies inside the grid, the unchecked fast path),
    20 p' sweeps (five launches of four: an odd count, started from the partner buffer) and the
    refinement, whose correction is added to p' by its last launch; 12 outer iterations."""
    cfg = CASES["cavity2d"]()
    cfg.nx, cfg.ny = shape
    cfg.sweep_block = block
    cfg.tile = tile
    cfg.p_refine = 1
    _log(f"[interior] block={block} tile={tile} grid={shape}")
    g = Solver(cfg)
    o = Oracle(cfg)
    rg = g.solve(12)
    ro = o.solve(12)
    _compare_all(g, o, cfg, f"interior block {block} tile {tile}")
    assert [tuple(r.sweeps) for r in rg["history"]] == [tuple(r.sweeps) for r in ro["history"]]


@pytest.mark.parametrize("shape", [(150, 20, 70), (70, 20, 40), (9, 17, 33)])
def test_3d_zmarch_multiple_tiles_and_chunks(shape):
    """The 3D fused sweep with several x/y tiles and z chunks (chunk edges recompute red planes;
    the tile is 128 x 8 cells over 32 planes), u, v, w with different sweep counts (launches of
    three, two and one equations)."""
    cfg = CASES["duct3d"]()
    cfg.nx, cfg.ny, cfg.nz = shape
    cfg.n_iter[0], cfg.n_iter[1], cfg.n_iter[2], cfg.n_iter[3] = 3, 2, 5, 7
    _log(f"[3d zmarch] grid={shape}")
    g = Solver(cfg)
    o = Oracle(cfg)
    rg = g.solve(3)
    ro = o.solve(3)
    _compare_all(g, o, cfg, f"3d {shape}")
    assert [tuple(r.sweeps) for r in rg["history"]] == [tuple(r.sweeps) for r in ro["history"]]


@pytest.mark.parametrize("p_refine", [0, 1])
@pytest.mark.parametrize("shape,n_p", [((150, 20, 70), 7), ((9, 17, 33), 8), ((40, 9, 5), 5), ((64, 16, 96), 20)])
def test_3d_two_pcorr_sweeps_per_launch(shape, n_p, p_refine):
    """sweep_block 2 in 3D: two red-black p' sweeps per launch (k_rb_zmarch3d_p2: 4-cell halo, the
    column's phases starting up to 3 planes below and ending up to 3 above it) on ragged x/y tiles
    and z chunks; an odd sweep count ends with a one-sweep launch; the refinement's sweeps too."""
    cfg = CASES["duct3d"]()
    cfg.nx, cfg.ny, cfg.nz = shape
    cfg.n_iter[3] = n_p
    cfg.sweep_block = 2
    cfg.p_refine = p_refine
    _log(f"[3d p' block 2] grid={shape} n_p={n_p} p_refine={p_refine}")
    g = Solver(cfg)
    o = Oracle(cfg)
    rg = g.solve(3)
    ro = o.solve(3)
    _compare_all(g, o, cfg, f"3d p' block 2 {shape}")
    assert [tuple(r.sweeps) for r in rg["history"]] == [tuple(r.sweeps) for r in ro["history"]]


def test_3d_two_pcorr_sweeps_stop_inside_a_block():
    """A loose p' tolerance stops the call after a sweep inside a two-sweep launch: the call ends
    at the end of that launch on both sides (the oracle's blocked rule), over 6 outer iterations
    of the 3D cavity."""
    cfg = CASES["cavity3d"]()
    cfg.nx, cfg.ny, cfg.nz = 70, 20, 40
    cfg.sweep_block = 2
    cfg.n_iter[3] = 20
    cfg.tol[3] = 0.5
    _log(f"[3d p' block 2 stop] grid={cfg.nx}x{cfg.ny}x{cfg.nz}")
    g = Solver(cfg)
    o = Oracle(cfg)
    rg = g.solve(6)
    ro = o.solve(6)
    _compare_all(g, o, cfg, "3d p' block 2 stop")
    sweeps = [tuple(r.sweeps) for r in rg["history"]]
    assert sweeps == [tuple(r.sweeps) for r in ro["history"]]
    assert any(sw[3] < cfg.n_iter[3] for sw in sweeps), sweeps
    assert all(sw[3] % 2 == 0 for sw in sweeps), sweeps


def test_3d_jacobi_parity():
    cfg = CASES["cavity3d"]()
    cfg.smoother = _abi.SMOOTHER_JACOBI
    cfg.n_iter[3] = 5
    g = Solver(cfg)
    o = Oracle(cfg)
    g.solve(3)
    o.solve(3)
    _compare_all(g, o, cfg, "3d jacobi")
