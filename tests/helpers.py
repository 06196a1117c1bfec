"""Shared helpers of the test suite: small cases, random states and bitwise comparison."""
import numpy as np

from incompflow import _abi

STATE_FIELDS = (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF,
                _abi.FIELD_WF, _abi.FIELD_PCORR)


def channel(nx=24, ny=12, iters=5):
    c = _abi.channel_config(nx, ny, iters)
    return c


def cavity(n=20, m=None, re=100.0, p_refine=0):
    c = _abi.cavity_config(n=n, re=re, p_refine=p_refine)
    if m is not None:
        c.ny = m
    c.stop_rule = _abi.STOP_NONE
    return c


def cavity3d(nx=10, ny=8, nz=6, re=100.0):
    c = _abi.cavity3d_config(n=nx, re=re)
    c.ny, c.nz = ny, nz
    c.stop_rule = _abi.STOP_NONE
    return c


def duct3d(nx=12, ny=6, nz=5):
    """3D channel: west inlet, east outlet, four walls, moving top wall."""
    c = cavity3d(nx, ny, nz)
    c.xmin, c.xmax = 0.0, 2.0
    c.bc_type[_abi.WEST] = _abi.INLET
    c.bc_type[_abi.EAST] = _abi.OUTLET
    c.bc_val[_abi.WEST][0] = 0.5
    c.bc_val[_abi.TOP][0] = 0.2
    return c


CASES = {
    "channel2d": lambda: channel(),
    "cavity2d": lambda: cavity(20, 16),
    "cavity3d": lambda: cavity3d(),
    "duct3d": lambda: duct3d(),
}


def random_state(cfg, seed):
    """Random but physically scaled fields (velocities O(1), pressure O(1))."""
    rng = np.random.default_rng(seed)
    st = {}
    for f in STATE_FIELDS:
        n = cfg.field_size(f)
        st[f] = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    if cfg.dim == 2:
        st[_abi.FIELD_W] = np.zeros(cfg.field_size(_abi.FIELD_W), np.float32)
        st[_abi.FIELD_WF] = np.zeros(cfg.field_size(_abi.FIELD_WF), np.float32)
    return st


def load_state(obj, st):
    for f, a in st.items():
        obj.set(f, a)


def bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32)).view(np.uint32)


def assert_bitwise(a, b, what=""):
    a = np.asarray(a, dtype=np.float32).ravel()
    b = np.asarray(b, dtype=np.float32).ravel()
    assert a.shape == b.shape, what
    diff = np.nonzero(bits(a) != bits(b))[0]
    if diff.size:
        i = diff[0]
        raise AssertionError(f"{what}: {diff.size} of {a.size} values differ; first at {i}: {a[i]!r} vs {b[i]!r}")


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))
