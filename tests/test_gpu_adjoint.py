"""The device adjoint: bit-identical to the FP32 oracle's adjoint iterations, and gradients that
agree with central differences of converged FP64 oracle solves (the reference algorithm)."""
import sys

import numpy as np
import pytest

from helpers import cavity, duct3d
from incompflow import Solver, _abi
from oracle.oracle import Oracle
from test_oracle_adjoint import channel_case, converged, layout, pressure_drop

pytestmark = pytest.mark.gpu

LAMBDA_FIELDS = (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF, _abi.FIELD_WF)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case(name):
    prof = None
    if name == "channel":
        c, prof = channel_case(24, 10, noise=1e-2)
    elif name == "cavity":
        c = cavity(20, 16, re=400.0)
    elif name == "duct3d":
        c = duct3d(10, 6, 5)
    else:                                   # inlet west, outlets east and north, moving south wall
        c = cavity(18, 12, re=200.0)
        c.bc_type[_abi.WEST] = _abi.INLET
        c.bc_val[_abi.WEST][0] = 0.5
        c.bc_type[_abi.EAST] = _abi.OUTLET
        c.bc_type[_abi.NORTH] = _abi.OUTLET
        c.bc_val[_abi.NORTH][0] = 0.0
        c.bc_val[_abi.SOUTH][0] = 0.3
    c.stop_rule = _abi.STOP_NONE
    return c, prof


def _split(c, flat):
    N = c.nx * c.ny * c.nz
    sizes = [N, N, N, N, (c.nx + 1) * c.ny * c.nz, c.nx * (c.ny + 1) * c.nz, c.nx * c.ny * (c.nz + 1)]
    out, off = {}, 0
    for f, n in zip(LAMBDA_FIELDS, sizes):
        out[f] = flat[off:off + n]
        off += n
    return out


@pytest.mark.parametrize("name", ["channel", "cavity", "duct3d", "outlet_north"])
def test_adjoint_iterations_equal_the_oracle_bit_for_bit(name):
    c, prof = _case(name)
    _log(f"[adjoint] {name}: forward 30 iterations, then 9 adjoint iterations")
    s = Solver(c)
    o = Oracle(c)
    if prof is not None:
        s.set_face_profile(_abi.WEST, prof)
        o.set_face_profile(_abi.WEST, prof)
    s.solve(30, history=False)
    o.solve(30, history=False)
    for f in (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF):
        assert np.array_equal(s.get(f, shaped=False), o.get(f, shaped=False))
    inlet = c.bc_type[_abi.WEST] == _abi.INLET
    r = s.adjoint(_abi.OBJ_PRESSURE_DROP, _abi.WEST, max_iter=9, tol=-1.0, profile_gradient=inlet)
    assert r["iterations"] == 9
    o.adj_setup(0)
    o.adj_iterate(9)
    want = _split(c, o.adj_lambda())
    for f in LAMBDA_FIELDS:
        if c.dim == 2 and f in (_abi.FIELD_W, _abi.FIELD_WF):
            continue
        got = s.adjoint_lambda(f, shaped=False)
        w = want[f].astype(np.float32)
        assert np.array_equal(got.view(np.uint32), w.view(np.uint32)), (name, f, np.abs(got - w).max())
    g = o.adj_gradient()
    assert abs(r["dJ_dmu"] - g[-1]) <= 1e-9 * abs(g[-1]), (r["dJ_dmu"], g[-1])
    if inlet:
        _, _, poff = layout(c)
        gw = g[poff[_abi.WEST]:poff[_abi.WEST + 1]]
        assert np.array_equal(r["dJ_dprofile"], gw)
    s.close()


_WIDE = {}


def _wide_oracle():
    """300x100 channel (interior tiles for every 2D tile shape), u with 7 sweeps and p' with 13
    (launches with fewer sweeps than the tile's K, odd and even launch counts)."""
    if not _WIDE:
        c, prof = channel_case(300, 100, noise=1e-2)
        c.stop_rule = _abi.STOP_NONE
        c.n_iter[0] = 7
        c.n_iter[3] = 13
        o = Oracle(c)
        o.set_face_profile(_abi.WEST, prof)
        o.solve(30, history=False)
        o.adj_setup(0)
        o.adj_iterate(9)
        _WIDE.update(cfg=c, prof=prof, lam=_split(c, o.adj_lambda()), grad=o.adj_gradient())
    return _WIDE


@pytest.mark.parametrize("tile", range(7))
def test_adjoint_tiles_equal_the_oracle_on_interior_tiles(tile):
    w = _wide_oracle()
    c = w["cfg"]
    _log(f"[adjoint] 300x100 channel, tile {tile}: forward 30 iterations, then 9 adjoint iterations")
    with Solver(c) as s:
        s.set_face_profile(_abi.WEST, w["prof"])
        s.solve(30, history=False)
        r = s.adjoint(_abi.OBJ_PRESSURE_DROP, _abi.WEST, max_iter=9, tol=-1.0, tile=tile)
        assert r["iterations"] == 9
        for f in LAMBDA_FIELDS:
            if f in (_abi.FIELD_W, _abi.FIELD_WF):
                continue
            got = s.adjoint_lambda(f, shaped=False)
            want = w["lam"][f].astype(np.float32)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (tile, f, np.abs(got - want).max())
    g = w["grad"]
    assert abs(r["dJ_dmu"] - g[-1]) <= 1e-9 * abs(g[-1]), (r["dJ_dmu"], g[-1])
    _, _, poff = layout(c)
    assert np.array_equal(r["dJ_dprofile"], g[poff[_abi.WEST]:poff[_abi.WEST + 1]])


def test_gradient_matches_differences_of_converged_fp64_solves():
    # FP32 stalls at a scaled residual of ~4e-5 on this grid (reached by iteration ~3000): run
    # the forward past it, and the adjoint to its own FP32 floor
    c, prof = channel_case(16, 8, noise=1e-2)
    c.stop_rule = _abi.STOP_NONE
    c.num_outer_iter = 5000
    _log("[adjoint] 16x8 channel: FP32 forward 5000 iterations, adjoint until dJ/dmu settles")
    with Solver(c) as s:
        s.set_face_profile(_abi.WEST, prof)
        s.solve(0, history=False)
        r = s.adjoint(_abi.OBJ_PRESSURE_DROP, _abi.WEST, tol=1e-7, grad_tol=1e-6, check_every=500, max_iter=40000)
    assert r["converged"]
    h = 1e-4 * c.viscosity
    fd_mu = (pressure_drop(converged(_mu(c, c.viscosity + h), prof), c)
             - pressure_drop(converged(_mu(c, c.viscosity - h), prof), c)) / (2 * h)
    assert abs(r["dJ_dmu"] - fd_mu) <= 1e-3 * abs(fd_mu), (r["dJ_dmu"], fd_mu)
    gw = r["dJ_dprofile"]
    for j in (0, c.ny // 2, c.ny - 1):
        hp = 1e-4 * 0.02
        pp, pm = prof.copy(), prof.copy()
        pp[j] += hp
        pm[j] -= hp
        fd = (pressure_drop(converged(c, pp), c) - pressure_drop(converged(c, pm), c)) / (2 * hp)
        assert abs(gw[j] - fd) <= 1e-3 * np.abs(gw).max(), (j, gw[j], fd)


def _mu(c, mu):
    c2 = c.copy()
    c2.viscosity = mu
    return c2


def test_torch_binding_on_the_device():
    torch = pytest.importorskip("torch")
    from incompflow.torch_adjoint import HipBackend, steady_objective
    c, prof = channel_case(16, 8, noise=1e-2)
    c.stop_rule = _abi.STOP_NONE
    c.num_outer_iter = 5000
    be = HipBackend(c, adjoint_tol=1e-6, adjoint_max_iter=20000, require_converged=False)
    runs = []
    solve = be.solve
    be.solve = lambda pr, m: runs.append(solve(pr, m)) or runs[-1]
    p = torch.tensor(prof, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(c.viscosity, dtype=torch.float64, requires_grad=True)
    J = steady_objective(p, mu, be)
    J.backward()
    rep = runs[-1].adjoint_report
    assert np.array_equal(p.grad.numpy(), rep["dJ_dprofile"])
    assert mu.grad.item() == rep["dJ_dmu"]
    # the objective the binding returns is the FP64 oracle's to FP32 accuracy
    J64 = pressure_drop(converged(c, prof), c)
    assert abs(J.item() - J64) <= 1e-4 * abs(J64)
    runs[-1].close()
    # a solve that runs out of iterations without meeting its stop rule is refused by default
    c2 = c.copy()
    c2.stop_rule = _abi.STOP_SCALED
    c2.tol_residual = 1e-12
    c2.num_outer_iter = 50
    with pytest.raises(RuntimeError, match="did not meet its stop rule"):
        HipBackend(c2).solve(prof, c.viscosity)


def _fp64_mg(c, mu, prof, tol=1e-10):
    """Converged FP64 oracle solve with the multigrid p' solve (the converged state does not depend
    on the p' solver), returning the pressure drop."""
    c2 = c.copy()
    c2.viscosity = mu
    c2.p_solver, c2.p_refine, c2.mg_over = _abi.PSOLVE_MG, 0, 1.0
    c2.stop_rule, c2.tol_residual, c2.num_outer_iter = _abi.STOP_SCALED, tol, 200000
    o = Oracle(c2, fp64=True)
    o.set_face_profile(_abi.WEST, prof)
    r = o.solve(200000, history=False)
    assert r["converged"], "FP64 solve did not converge"
    J = pressure_drop(o, c2)
    o.close()
    return J


def test_gradient_at_a_reduced_adjoint_channel_matches_fp64_differences():
    """The adjoint channel's physics on a 128x32 grid (the 1024x256 case reduced 8x): the device's
    FP32 adjoint gradient along mu, along the profile itself and for three single inlet faces agrees
    with central differences (step 1e-4 relative) of FP64 oracle solves converged to 1e-10, within
    1e-3 relative (the gradient's FP32 accuracy: the forward state stalls near the FP32 floor)."""
    c = _abi.adjoint_channel_config(128, 32)
    prof = np.array(_abi.adjoint_inlet_profile(c))
    c.stop_rule = _abi.STOP_NONE
    c.num_outer_iter = 20000
    _log("[adjoint] 128x32 channel: FP32 forward 20000 iterations, adjoint until dJ/dmu and dJ/dprofile settle")
    with Solver(c) as s:
        s.set_face_profile(_abi.WEST, prof)
        s.solve(0, history=False)
        r = s.adjoint(_abi.OBJ_PRESSURE_DROP, _abi.WEST, tol=1e-8, grad_tol=1e-6, check_every=500, max_iter=100000)
    assert r["converged"] and r["grad_change"] <= 1e-6
    mu, h = c.viscosity, 1e-4
    fd_mu = (_fp64_mg(c, mu * (1 + h), prof) - _fp64_mg(c, mu * (1 - h), prof)) / (2 * h * mu)
    assert abs(r["dJ_dmu"] - fd_mu) <= 1e-3 * abs(fd_mu), (r["dJ_dmu"], fd_mu)
    g = r["dJ_dprofile"]
    fd_dir = (_fp64_mg(c, mu, prof * (1 + h)) - _fp64_mg(c, mu, prof * (1 - h))) / (2 * h)
    assert abs(float(np.dot(prof, g)) - fd_dir) <= 1e-3 * abs(fd_dir), (float(np.dot(prof, g)), fd_dir)
    hp = 1e-4 * 0.02
    for j in (0, c.ny // 3, c.ny - 1):
        pp, pm = prof.copy(), prof.copy()
        pp[j] += hp
        pm[j] -= hp
        fd = (_fp64_mg(c, mu, pp) - _fp64_mg(c, mu, pm)) / (2 * hp)
        assert abs(g[j] - fd) <= 1e-3 * np.abs(g).max(), (j, g[j], fd)
