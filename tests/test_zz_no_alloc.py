"""ifl_solve_noalloc (the solve that allocates no device memory) against ifl_create +
ifl_set_field + ifl_solve from the same fields: fields, iteration count, convergence flag and
records bit for bit.  Calls follow one another in one process, larger cases before smaller ones,
so each call runs on scratch that the previous call left dirty."""
import sys

import numpy as np
import pytest
import torch

from helpers import assert_bitwise, cavity, cavity3d, channel, duct3d
from incompflow import IflError, Solver, _abi, solve_noalloc

pytestmark = pytest.mark.gpu

A = _abi


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fields(cfg):
    f = [A.FIELD_U, A.FIELD_V, A.FIELD_P]
    if cfg.dim == 3:
        f.append(A.FIELD_W)
    if cfg.energy:
        f.append(A.FIELD_TEMP)
    return f


def _initial(cfg, seed):
    rng = np.random.default_rng(seed)
    n = cfg.nx * cfg.ny * cfg.nz
    init = {f: (0.05 * rng.standard_normal(n)).astype(np.float32) for f in _fields(cfg)}
    if cfg.energy:
        init[A.FIELD_TEMP] = (300.0 + 5.0 * rng.standard_normal(n)).astype(np.float32)
    return init


def _profile(cfg):
    return (0.4 + 0.1 * np.sin(np.arange(cfg.ny * cfg.nz))).astype(np.float32)


def _by_handle(cfg, init, prof, max_iter):
    with Solver(cfg) as s:
        for f, a in init.items():
            s.set(f, a)
        if prof is not None:
            s.set_face_profile(A.WEST, prof.astype(np.float64))
        r = s.solve(max_iter, history=True)
        return r, {f: s.get(f, shaped=False) for f in init}


def _by_noalloc(cfg, init, prof, max_iter, stream=None):
    t = {f: torch.from_numpy(a.copy()).cuda() for f, a in init.items()}
    fp = {A.WEST: torch.from_numpy(prof).cuda()} if prof is not None else None
    if stream is not None:
        stream.wait_stream(torch.cuda.current_stream())
    r = solve_noalloc(cfg, t[A.FIELD_U], t[A.FIELD_V], t[A.FIELD_P], w=t.get(A.FIELD_W), temp=t.get(A.FIELD_TEMP),
                      face_profiles=fp, max_iter=max_iter, stream=stream, history=True)
    if stream is not None:
        torch.cuda.current_stream().wait_stream(stream)
    torch.cuda.synchronize()
    return r, {f: v.cpu().numpy() for f, v in t.items()}


def _compare(label, cfg, max_iter, prof=None, seed=3, stream=None):
    _log(f"[noalloc] {label}: {cfg.nx}x{cfg.ny}x{cfg.nz}, {max_iter} iterations")
    init = _initial(cfg, seed)
    rh, fh = _by_handle(cfg, init, prof, max_iter)
    rn, fn = _by_noalloc(cfg, init, prof, max_iter, stream)
    assert rn["iterations"] == rh["iterations"], label
    assert rn["converged"] == rh["converged"] and rn["status"] == rh["status"], label
    assert rn["final_scaled"] == rh["final_scaled"], label
    assert rn["sweeps_total"] == rh["sweeps_total"], label
    assert len(rn["history"]) == len(rh["history"])
    for a, b in zip(rn["history"], rh["history"]):
        assert a.scaled == b.scaled and a.mass_norm == b.mass_norm and list(a.sweeps) == list(b.sweeps), label
    for f in fh:
        assert_bitwise(fn[f], fh[f], f"{label}: {A.FIELD_NAMES[f]}")
    return rh


def test_two_calls_in_a_row_the_second_smaller():
    big = cavity(200, 130, p_refine=1)
    big.sweep_block = 4
    big.stop_rule = A.STOP_SCALED
    big.tol_residual = 2e-3
    r = _compare("cavity 200x130, scaled stop", big, 600)
    assert r["converged"] and r["iterations"] < 600
    small = channel(64, 24, 60)
    small.p_solver, small.p_refine = A.PSOLVE_MG, 0
    _compare("channel 64x24, profile, multigrid", small, 60, prof=_profile(small))


def test_3d_energy_jacobi_and_a_side_stream():
    _compare("duct 24x12x10", duct3d(24, 12, 10), 12)
    mg3 = duct3d(36, 20, 18)
    mg3.p_solver, mg3.p_refine, mg3.mg_over = A.PSOLVE_MG, 0, 1.4
    _compare("duct 36x20x18, multigrid", mg3, 10)
    two = duct3d(36, 20, 18)
    two.sweep_block = 2
    _compare("duct 36x20x18, two p' sweeps per launch", two, 8)
    _compare("cavity3d 20x17x33 on a side stream", cavity3d(20, 17, 33, re=300.0), 6, stream=torch.cuda.Stream())
    hot = cavity(48, 40)
    hot.energy = 1
    _compare("cavity 48x40 with temperature", hot, 20)
    jac = cavity(40, 24)
    jac.smoother = A.SMOOTHER_JACOBI
    _compare("cavity 40x24, Jacobi", jac, 15)


def test_a_case_beyond_the_scratch_is_refused():
    c = A.cavity_config(n=800)
    t = [torch.zeros(800 * 800, device="cuda") for _ in range(3)]
    with pytest.raises(IflError) as e:
        solve_noalloc(c, t[0], t[1], t[2], max_iter=1)
    assert e.value.status == A.IFL_ERR_ARG and "bytes of device scratch" in str(e.value)
    small = cavity(32, 32)
    _compare("cavity 32x32 after the refusal", small, 10)
