"""Host-side checks of the C-ABI library (no GPU needed: nothing here launches a kernel)."""
import ctypes as C
import os
import re

import pytest

from incompflow import _abi
from incompflow import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "incompflow.h")


def declared_entry_points():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ifl_[a-z_]+)\s*\(", text)) - {"ifl_config", "ifl_report"})


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_entry_points()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/incompflow.h but not exported"
    assert {n for n, _, _ in _lib.EXPORTS} == set(names)


def test_struct_layouts_match_header():
    # field order/sizes: the library itself fills a default config through the ABI struct
    c = _abi.Config()
    _lib.lib().ifl_config_default(C.byref(c))
    assert (c.dim, c.nx, c.ny, c.nz) == (2, 1024, 256, 1)          # include/config.h:6-10
    assert (c.xmin, c.xmax, c.ymin, c.ymax) == (-2.0, 2.0, -0.5, 0.5)
    assert (c.density, c.viscosity) == (50.0, 1.0)
    assert list(c.bc_type) == [_abi.OUTLET, _abi.INLET, _abi.WALL, _abi.WALL, _abi.WALL, _abi.WALL]
    assert c.bc_val[_abi.WEST][0] == 1.0
    assert list(c.n_iter) == [10, 10, 10, 20] and list(c.relax) == [0.75, 0.75, 0.75, 0.25]
    assert list(c.tol) == [1e-5, 1e-5, 1e-5, 1e-3] and c.tol_mass == 1e-9
    assert c.init[4] == 273.0 and c.stop_rule == _abi.STOP_FAITHFUL and c.log_every == 1000


@pytest.mark.parametrize("mutate,msg", [
    (lambda c: setattr(c, "nx", 3), "nx"),
    (lambda c: setattr(c, "dim", 4), "dim"),
    (lambda c: setattr(c, "nz", 2), "nz"),
    (lambda c: c.bc_type.__setitem__(0, 7), "bc_type"),
    (lambda c: c.n_iter.__setitem__(3, 0), "n_iter"),
    (lambda c: setattr(c, "smoother", _abi.SMOOTHER_LEX), "smoother"),
    (lambda c: setattr(c, "xmax", c.xmin), "domain"),
    (lambda c: setattr(c, "tile", 9), "tile"),
    (lambda c: setattr(c, "p_solver", 5), "p_solver"),
    (lambda c: (setattr(c, "p_solver", _abi.PSOLVE_MG), setattr(c, "p_refine", 1)), "p_refine"),
    (lambda c: (setattr(c, "p_solver", _abi.PSOLVE_MG), setattr(c, "nx", 4), setattr(c, "ny", 60000)), "elongated"),
    (lambda c: setattr(c, "mg_over", 2.5), "mg_over"),
    (lambda c: setattr(c, "mg_sweeps", 4), "mg_sweeps"),
    (lambda c: setattr(c, "mg_cycles", 9), "mg_cycles"),
    (lambda c: setattr(c, "mg_coarse_sweeps", -1), "mg_coarse_sweeps"),
    (lambda c: setattr(c, "fail_if_not_converged", 2), "fail_if_not_converged"),
])
def test_invalid_config_rejected_before_any_device_work(mutate, msg):
    c = _abi.default_config()
    mutate(c)
    h = C.c_void_p()
    st = _lib.lib().ifl_create(C.byref(c), C.byref(h))
    assert st == _abi.IFL_ERR_ARG
    assert not h.value
    assert msg in _lib.lib().ifl_last_error().decode()


def test_null_arguments():
    L = _lib.lib()
    assert L.ifl_create(None, None) == _abi.IFL_ERR_ARG
    assert L.ifl_destroy(None) == _abi.IFL_OK
    assert L.ifl_solve(None, 1, None, None, 0) == _abi.IFL_ERR_ARG
    assert L.ifl_batch_destroy(None) == _abi.IFL_OK
    assert L.ifl_batch_solve(None, 1, None) == _abi.IFL_ERR_ARG
    assert L.ifl_batch_size(None) == 0
    h = C.c_void_p()
    assert L.ifl_batch_case(None, 0, C.byref(h)) == _abi.IFL_ERR_ARG


@pytest.mark.parametrize("mutate,msg", [
    (lambda c: setattr(c, "smoother", _abi.SMOOTHER_JACOBI), "must agree"),
    (lambda c: c.n_iter.__setitem__(3, 7), "must agree"),
    (lambda c: setattr(c, "p_refine", 2), "must agree"),
    (lambda c: setattr(c, "sweep_block", 2), "must agree"),
    (lambda c: setattr(c, "tile", 3), "must agree"),
    (lambda c: setattr(c, "nx", 2), "case 1: nx and ny"),
    (lambda c: (setattr(c, "p_solver", _abi.PSOLVE_MG), setattr(c, "p_refine", 0)), "p' sweeps"),
])
def test_batch_rejects_mismatched_cases_before_any_device_work(mutate, msg):
    cs = (_abi.Config * 2)(_abi.cavity_config(n=32), _abi.cavity_config(n=48))
    mutate(cs[1])
    h = C.c_void_p()
    st = _lib.lib().ifl_batch_create(cs, 2, C.byref(h))
    assert st == _abi.IFL_ERR_ARG
    assert not h.value
    assert msg in _lib.lib().ifl_last_error().decode()


def test_batch_workload_generator():
    """The seeded uneven batch: 64 cases, sizes in 16*{4..32}, Re in [10, 1000], both types."""
    cs = _abi.uneven_batch_configs()
    assert len(cs) == 64
    sizes = [(c.nx, c.ny) for c in cs]
    assert all(n % 16 == 0 and 64 <= n <= 512 for s in sizes for n in s)
    res = [1.0 / c.viscosity for c in cs]
    assert all(10.0 <= r <= 1000.0 for r in res)
    kinds = [c.bc_type[_abi.WEST] for c in cs]
    assert 16 < kinds.count(_abi.INLET) < 48
    assert sum(c.nx * c.ny for c in cs) == 5340160      # fixed by the seed
    # SplitMix64 known answer (seed 0): first output 0xe220a8397b1dcdaf
    g = _abi.splitmix64(0)
    assert next(g) == (0xE220A8397B1DCDAF >> 11) / float(1 << 53)
