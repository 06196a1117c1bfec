"""Binary state-file format (include/incompflow.h, ifl_state_header): host-side checks, no GPU.

The library writes and reads the format in C++ (ifl_save_state / ifl_load_state); the package
mirrors it in numpy (read_state / write_state) so that state files can be inspected or produced
without a GPU.  The GPU tests (test_gpu_state.py) check that both sides agree.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from incompflow import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fnv1a64_known_answers():
    # published FNV-1a 64-bit test vectors
    assert _abi.fnv1a64(b"") == 0xcbf29ce484222325
    assert _abi.fnv1a64(b"a") == 0xaf63dc4c8601ec8c
    assert _abi.fnv1a64(b"foobar") == 0x85944171f73967e8


def test_header_mirror_matches_c_header():
    text = open(os.path.join(ROOT, "include", "incompflow.h")).read()
    body = text[text.index("typedef struct ifl_state_header"):text.index("} ifl_state_header;")]
    names = []
    for decl in re.findall(r"^\s+\w+\s+([\w\[\], ]+);", body, flags=re.M):
        names += [re.sub(r"\[\d+\]", "", d).strip() for d in decl.split(",")]
    assert names == [n for n, _ in _abi.StateHeader._fields_]
    assert C.sizeof(_abi.StateHeader) == 160      # static_assert in the library
    assert '#define IFL_STATE_MAGIC "IFLSTATE"' in text and "#define IFL_STATE_VERSION 1" in text


@pytest.mark.parametrize("dims", [(2, 7, 5, 1), (3, 4, 6, 5)])
def test_round_trip(tmp_path, dims):
    dim, nx, ny, nz = dims
    cfg = _abi.Config()
    cfg.dim, cfg.nx, cfg.ny, cfg.nz = dims
    rng = np.random.default_rng(7)
    fields = {f: rng.standard_normal(cfg.field_size(f)).astype(np.float32) for f in _abi.STATE_FIELDS}
    path = tmp_path / "s.bin"
    _abi.write_state(path, dim, nx, ny, nz, fields, iter=123, norm_max=list(range(1, 8)),
                     res_max=[0.5, 0.25, 0.125, 0.0625, 0.03125], last_scaled=3e-4, last_converged=1)
    size = os.path.getsize(path)
    assert size == 160 + 4 * sum(cfg.field_size(f) for f in _abi.STATE_FIELDS)
    hd, got = _abi.read_state(path)
    assert (hd.dim, hd.nx, hd.ny, hd.nz, hd.iter, hd.last_converged) == (dim, nx, ny, nz, 123, 1)
    assert list(hd.norm_max) == list(range(1, 8)) and list(hd.res_max) == [0.5, 0.25, 0.125, 0.0625, 0.03125]
    for f in _abi.STATE_FIELDS:
        assert got[f].shape == cfg.shape(f)
        assert np.array_equal(got[f].ravel().view(np.uint32), fields[f].view(np.uint32))
    # field order on disk: u first, wf last
    raw = open(path, "rb").read()
    assert raw[160:164] == fields[_abi.FIELD_U][:1].tobytes()
    assert raw[-4:] == fields[_abi.FIELD_WF][-1:].tobytes()


def test_corruption_detected(tmp_path):
    path = tmp_path / "s.bin"
    _abi.write_state(path, 2, 4, 4, 1, {})
    raw = bytearray(open(path, "rb").read())
    bad = tmp_path / "bad.bin"
    raw[200] ^= 1
    bad.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        _abi.read_state(bad)
    bad.write_bytes(bytes(raw[:-4]))
    with pytest.raises(ValueError, match="payload size"):
        _abi.read_state(bad)
    bad.write_bytes(b"NOTSTATE" + bytes(raw[8:]))
    with pytest.raises(ValueError, match="not a version"):
        _abi.read_state(bad)
    with pytest.raises(ValueError, match="values"):
        _abi.write_state(path, 2, 4, 4, 1, {_abi.FIELD_U: np.zeros(3)})
