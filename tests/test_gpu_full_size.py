"""The benchmark configurations at their full sizes, on the device against the FP32 oracle.

The other GPU tests compare reduced grids bit for bit; these run the production shapes for as
many outer iterations as the host oracle finishes in seconds:
  * the 1024^2 lid-driven cavity, precision form (FP64-defect refinement, the bench's launch
    settings: 128x32 tiles, five sweeps per launch);
  * the 256^3 cavity (Re 400) with the reference's p' sweeps and with the multigrid p' solve;
  * the 1024x256 adjoint channel with its seeded noisy inlet: forward iterations, then adjoint
    iterations and the gradient;
  * all 64 cases of the seeded uneven batch in one batch: every case equal to its own solve,
    three of them also to the oracle.
"""
import sys

import numpy as np
import pytest

from helpers import assert_bitwise
from incompflow import Batch, Solver, _abi
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FIELDS_2D = (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF)
FIELDS_3D = FIELDS_2D + (_abi.FIELD_W, _abi.FIELD_WF)
LAMBDA_FIELDS = (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _records(h):
    keys = ("p_norm", "u_norm", "v_norm", "w_norm", "mass_norm", "res_cont", "res_u", "res_v", "res_w", "scaled")
    return np.array([[getattr(r, k) for k in keys] for r in h])


def _same(g, o, cfg, fields, what):
    for f in fields:
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), f"{what}: {_abi.FIELD_NAMES[f]}")


def test_c2_cavity_1024_precision_form():
    cfg = _abi.cavity_config(n=1024, p_refine=1, tol_residual=1e-4)
    cfg.tile, cfg.sweep_block, cfg.sweep_block_mom = _abi.TILE_128x32, 5, 5
    _log("[full size] cavity 1024x1024, p_refine 1, 20 outer iterations")
    g, o = Solver(cfg), Oracle(cfg)
    rg, ro = g.solve(20), o.solve(20)
    assert rg["iterations"] == ro["iterations"] == 20
    _same(g, o, cfg, FIELDS_2D, "cavity 1024^2")
    assert [tuple(r.sweeps) for r in rg["history"]] == [tuple(r.sweeps) for r in ro["history"]]
    np.testing.assert_allclose(_records(rg["history"]), _records(ro["history"]), rtol=1e-11, atol=1e-300)
    g.close()


@pytest.mark.parametrize("p_solver", ["sweeps", "multigrid"])
def test_c3_cavity_256_cubed(p_solver):
    cfg = _abi.cavity3d_config(n=256)
    cfg.p_refine = 0
    iters = 2
    if p_solver == "multigrid":
        cfg.p_solver, cfg.mg_cycles, cfg.mg_over = _abi.PSOLVE_MG, 1, 1.8
        iters = 1
    _log(f"[full size] cavity 256^3, {p_solver}, {iters} outer iterations")
    g, o = Solver(cfg), Oracle(cfg)
    rg, ro = g.solve(iters), o.solve(iters)
    assert rg["iterations"] == ro["iterations"] == iters
    _same(g, o, cfg, FIELDS_3D, f"cavity 256^3 {p_solver}")
    np.testing.assert_allclose(_records(rg["history"]), _records(ro["history"]), rtol=1e-11, atol=1e-300)
    g.close()
    o.close()


def test_c5_channel_1024x256_forward_and_adjoint():
    cfg = _abi.adjoint_channel_config(1024, 256)
    cfg.stop_rule = _abi.STOP_NONE
    cfg.sweep_block = 4
    prof = np.array(_abi.adjoint_inlet_profile(cfg))
    _log("[full size] adjoint channel 1024x256: 200 forward iterations, 20 adjoint iterations")
    g, o = Solver(cfg), Oracle(cfg)
    g.set_face_profile(_abi.WEST, prof)
    o.set_face_profile(_abi.WEST, prof)
    g.solve(200, history=False)
    o.solve(200, history=False)
    _same(g, o, cfg, FIELDS_2D, "channel 1024x256 forward")
    r = g.adjoint(_abi.OBJ_PRESSURE_DROP, _abi.WEST, max_iter=20, tol=-1.0, profile_gradient=True)
    assert r["iterations"] == 20
    o.adj_setup(0)
    o.adj_iterate(20)
    lam = o.adj_lambda()
    N = cfg.nx * cfg.ny
    sizes = {_abi.FIELD_U: N, _abi.FIELD_V: N, _abi.FIELD_W: N, _abi.FIELD_P: N,
             _abi.FIELD_UF: (cfg.nx + 1) * cfg.ny, _abi.FIELD_VF: cfg.nx * (cfg.ny + 1), _abi.FIELD_WF: 2 * N}
    off = 0
    for f in (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF, _abi.FIELD_WF):
        want = lam[off:off + sizes[f]].astype(np.float32)
        off += sizes[f]
        if f in LAMBDA_FIELDS:
            assert_bitwise(g.adjoint_lambda(f, shaped=False), want, f"adjoint {_abi.FIELD_NAMES[f]}")
    gr = o.adj_gradient()
    assert abs(r["dJ_dmu"] - gr[-1]) <= 1e-9 * abs(gr[-1])
    g.close()


def test_c4_full_uneven_batch():
    cfgs = _abi.uneven_batch_configs(tol_residual=1e-4, max_iter=60)
    for c in cfgs:
        c.tile, c.sweep_block, c.sweep_block_mom = _abi.TILE_64x16, 5, 0
        c.num_outer_iter = 60
    _log(f"[full size] uneven batch, {len(cfgs)} cases, 60 outer iterations")
    with Batch(cfgs) as b:
        st, reps = b.solve(60)
        got = [{f: b.cases[i].get(f, shaped=False) for f in FIELDS_2D} for i in range(len(cfgs))]
    assert st == 0
    for i, c in enumerate(cfgs):
        with Solver(c) as s:
            r = s.solve(60, history=False)
            assert reps[i]["iterations"] == r["iterations"], i
            for f in FIELDS_2D:
                assert_bitwise(got[i][f], s.get(f, shaped=False), f"batch case {i} {_abi.FIELD_NAMES[f]}")
    sizes = [c.nx * c.ny for c in cfgs]
    for i in (int(np.argmax(sizes)), 47, next(k for k, c in enumerate(cfgs) if c.bc_type[_abi.WEST] == _abi.INLET)):
        o = Oracle(cfgs[i])
        o.solve(60, history=False)
        for f in FIELDS_2D:
            assert_bitwise(got[i][f], o.get(f, shaped=False), f"batch case {i} vs oracle {_abi.FIELD_NAMES[f]}")
        o.close()


@pytest.mark.parametrize("shape", [(16384, 8192, 1), (512, 512, 512)])
def test_largest_grids(shape):
    """2^27 cells (half the largest grid the library accepts, 2^28): every index and byte offset of
    a plane stays far inside 31 bits; the device equals the oracle after the first iterations."""
    nx, ny, nz = shape
    if nz == 1:
        cfg = _abi.cavity_config(n=nx, p_refine=1, tol_residual=1e-4)
        cfg.nx, cfg.ny = nx, ny
        cfg.tile, cfg.sweep_block, cfg.sweep_block_mom = _abi.TILE_128x32, 5, 5
        iters, fields = 2, FIELDS_2D
    else:
        cfg = _abi.cavity3d_config(n=nx)
        cfg.p_refine = 0
        iters, fields = 1, FIELDS_3D
    _log(f"[full size] largest grid {nx}x{ny}x{nz}, {iters} outer iterations")
    g = Solver(cfg)
    rg = g.solve(iters)
    o = Oracle(cfg)
    ro = o.solve(iters)
    assert rg["iterations"] == ro["iterations"] == iters
    _same(g, o, cfg, fields, f"grid {nx}x{ny}x{nz}")
    np.testing.assert_allclose(_records(rg["history"]), _records(ro["history"]), rtol=1e-11, atol=1e-300)
    g.close()
    o.close()
