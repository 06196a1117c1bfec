"""Binary state files on the GPU: ifl_save_state / ifl_load_state (runs on an MI355X).

A state file holds the fields and the outer-iteration monitor, which is everything the next
outer iteration reads, so a solve split at any iteration and continued in a new handle must give
the uninterrupted solve's iterates, history and stop iteration bit for bit; the FP32 oracle
continued from the same fields must agree too.
"""
import os
import sys

import numpy as np
import pytest

from helpers import assert_bitwise, cavity, cavity3d, channel, duct3d
from incompflow import _abi
from incompflow import IflError, Solver, read_state, write_state
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FIELDS2 = (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF)
FIELDS3 = FIELDS2 + (_abi.FIELD_W, _abi.FIELD_WF)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scaled_cavity(n=24):
    c = cavity(n, n)
    c.stop_rule = _abi.STOP_SCALED
    c.tol_residual = 1e-3
    c.num_outer_iter = 5000
    return c


def _profiled_channel():
    c = channel(32, 12)
    c.stop_rule = _abi.STOP_FIXED
    return c


CASES = {
    "cavity2d_scaled": (_scaled_cavity, None),
    "channel2d_profile": (_profiled_channel, np.linspace(0.0005, 0.0015, 12)),
    "cavity3d": (lambda: cavity3d(), None),
    "duct3d": (lambda: duct3d(), None),
}


def _make(name):
    mk, prof = CASES[name]
    cfg = mk()
    g = Solver(cfg)
    if prof is not None:
        g.set_face_profile(_abi.WEST, prof)
    return cfg, g


def _records(r):
    keys = ("scaled", "res_cont", "res_u", "res_v", "mass_norm", "rel_p", "rel_u", "rel_uf")
    return [tuple(getattr(x, k) for k in keys) + (x.converged,) for x in r["history"]]


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("split", [0, 3, 9])
def test_resume_bitwise(tmp_path, name, split):
    """split 0: saved before the first iteration (faces not yet initialised); 3: inside the
    scaled-residual normaliser window (iterations 1..5); 9: after it."""
    _log(f"state resume {name} split {split}")
    more = 7
    cfg, a = _make(name)
    if split:
        a.solve(split, history=False)
    path = str(tmp_path / "s.bin")
    a.save_state(path)
    ra = a.solve(more, history=True)
    cfg2, b = _make(name)
    b.load_state(path)
    rb = b.solve(more, history=True)
    assert ra["iterations"] == rb["iterations"]
    assert _records(ra) == _records(rb)
    for f in (FIELDS3 if cfg.dim == 3 else FIELDS2):
        assert_bitwise(b.get(f, shaped=False), a.get(f, shaped=False), f"{name}: {_abi.FIELD_NAMES[f]}")
    a.close()
    b.close()


def test_resume_stops_at_the_same_iteration(tmp_path):
    _log("state resume: stop iteration")
    cfg = _scaled_cavity(20)
    a = Solver(cfg)
    ra = a.solve(0, history=False)
    assert ra["converged"]
    total = ra["iterations"]
    b = Solver(cfg)
    b.solve(total // 2, history=False)
    path = str(tmp_path / "half.bin")
    b.save_state(path)
    b.close()
    c = Solver(cfg)
    c.load_state(path)
    rc = c.solve(0, history=False)
    assert rc["converged"] and total // 2 + rc["iterations"] == total
    assert rc["final_scaled"] == ra["final_scaled"]
    assert_bitwise(c.get(_abi.FIELD_P, shaped=False), a.get(_abi.FIELD_P, shaped=False), "p")
    a.close()
    c.close()


@pytest.mark.parametrize("name", ["cavity2d_scaled", "duct3d"])
def test_device_file_matches_host_reader_and_oracle(tmp_path, name):
    """The device-written file parses with the numpy reader; the oracle continued from its fields
    matches the device continued from the file bit for bit."""
    _log(f"state vs oracle {name}")
    cfg = CASES[name][0]()
    cfg.stop_rule = _abi.STOP_NONE
    g = Solver(cfg)
    g.solve(6, history=False)
    path = str(tmp_path / "s.bin")
    g.save_state(path)
    hd, fields = read_state(path)
    assert (hd.dim, hd.nx, hd.ny, hd.nz, hd.iter) == (cfg.dim, cfg.nx, cfg.ny, cfg.nz, 6)
    for f, a in fields.items():
        assert_bitwise(a, g.get(f, shaped=False), f"reader {_abi.FIELD_NAMES[f]}")
    o = Oracle(cfg)
    for f, a in fields.items():
        o.set(f, a)
    o.set_iter(hd.iter)
    o.solve(4, history=False)
    g.solve(4, history=False)
    for f in (FIELDS3 if cfg.dim == 3 else FIELDS2):
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), f"{name}: {_abi.FIELD_NAMES[f]}")
    g.close()
    o.close()


def test_host_written_file_loads(tmp_path):
    """A file written by the numpy writer from the oracle's state continues on the device exactly
    as the oracle continues."""
    _log("state: host-written file")
    cfg = cavity(16, 12)
    o = Oracle(cfg)
    o.solve(5, history=False)
    flds = {f: o.get(f, shaped=False) for f in _abi.STATE_FIELDS}
    path = str(tmp_path / "host.bin")
    write_state(path, cfg.dim, cfg.nx, cfg.ny, cfg.nz, flds, iter=5)
    g = Solver(cfg)
    g.load_state(path)
    g.solve(3, history=False)
    o.solve(3, history=False)
    for f in FIELDS2:
        assert_bitwise(g.get(f, shaped=False), o.get(f, shaped=False), _abi.FIELD_NAMES[f])
    g.close()
    o.close()


def test_warm_start_other_viscosity(tmp_path):
    """Warm start of a case with 2% higher viscosity from a converged state: fewer iterations than
    from rest, and the same fixed point (FP32 floor) as the cold solve."""
    _log("state: warm start")
    cfg = _scaled_cavity(24)
    cfg.tol_residual = 1e-4
    a = Solver(cfg)
    assert a.solve(0, history=False)["converged"]
    path = str(tmp_path / "base.bin")
    a.save_state(path)
    a.close()
    c2 = cfg.copy()
    c2.viscosity = cfg.viscosity * 1.02
    cold = Solver(c2)
    rc = cold.solve(0, history=False)
    warm = Solver(c2)
    warm.load_state(path)
    rw = warm.solve(0, history=False)
    assert rc["converged"] and rw["converged"]
    assert rw["iterations"] < rc["iterations"]
    for f in (_abi.FIELD_U, _abi.FIELD_V):
        x, y = warm.get(f), cold.get(f)
        assert np.max(np.abs(x - y)) <= 2e-3 * np.max(np.abs(y)), _abi.FIELD_NAMES[f]
    cold.close()
    warm.close()


def test_load_errors_leave_the_handle_unchanged(tmp_path):
    _log("state: errors")
    cfg = cavity(16, 12)
    g = Solver(cfg)
    g.solve(4, history=False)
    good = str(tmp_path / "good.bin")
    g.save_state(good)
    before = g.get(_abi.FIELD_P, shaped=False).copy()
    other = Solver(cavity(16, 14))
    other_path = str(tmp_path / "other.bin")
    other.save_state(other_path)
    other.close()
    with pytest.raises(IflError) as e:
        g.load_state(other_path)
    assert e.value.status == _abi.IFL_ERR_ARG and "does not match" in str(e.value)
    raw = bytearray(open(good, "rb").read())
    cases = {"truncated": bytes(raw[:-8]), "flipped": bytes(raw[:300] + bytes([raw[300] ^ 0x40]) + raw[301:]),
             "magic": b"XFLSTATE" + bytes(raw[8:]), "longer": bytes(raw) + b"\0"}
    for what, data in cases.items():
        p = tmp_path / f"{what}.bin"
        p.write_bytes(data)
        with pytest.raises(IflError) as e:
            g.load_state(str(p))
        assert e.value.status == _abi.IFL_ERR_IO, what
    with pytest.raises(IflError) as e:
        g.load_state(str(tmp_path / "missing.bin"))
    assert e.value.status == _abi.IFL_ERR_IO
    assert_bitwise(g.get(_abi.FIELD_P, shaped=False), before, "p after failed loads")
    g.close()
