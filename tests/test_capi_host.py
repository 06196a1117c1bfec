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
