"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported only by tests, the smoke check and the CPU-baseline leg of bench.py, as the checker.
The library is built from oracle/simple_oracle.hpp + oracle/oracle_capi.cpp by oracle/Makefile.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
sys.path.insert(0, os.path.join(_ROOT, "incompflow-cuda_amd"))
from incompflow import _abi  # noqa: E402

# IFL_ORACLE_LIB selects another build of the same sources (the sanitizer build, oracle/_san/)
LIB_PATH = os.environ.get("IFL_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(_abi.Config), C.c_int, C.POINTER(vp)]
        L.orc_destroy.argtypes = [vp]
        L.orc_field_size.argtypes = [vp, C.c_int]
        L.orc_field_size.restype = C.c_size_t
        L.orc_get_field.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.c_size_t]
        L.orc_set_field.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.c_size_t]
        L.orc_set_face_profile.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.c_size_t]
        L.orc_run_stage.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.c_size_t]
        L.orc_solve.argtypes = [vp, C.c_int, C.POINTER(_abi.IterRecord), C.c_size_t,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_set_iter.argtypes = [vp, C.c_longlong]
        L.orc_last_error.restype = C.c_char_p
        L.orc_config_size.restype = C.c_size_t
        L.orc_record_size.restype = C.c_size_t
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_fault.argtypes = [C.c_int]
        ip = C.POINTER(C.c_int)
        L.orc_mg_build.argtypes = [vp, ip, ip]
        L.orc_mg_get.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_size_t]
        L.orc_mg_set.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_size_t]
        L.orc_mg_cycles.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        L.orc_mg_smooth.argtypes = [vp, C.c_int, C.c_int]
        dp = C.POINTER(C.c_double)
        L.orc_adj_setup.argtypes = [vp, C.c_int]
        L.orc_adj_sizes.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.orc_adj_set_weights.argtypes = [vp, dp]
        L.orc_adj_objective.argtypes = [vp]
        L.orc_adj_objective.restype = C.c_double
        L.orc_adj_tangent.argtypes = [vp, dp, dp, dp]
        L.orc_adj_transpose.argtypes = [vp, dp, dp, dp]
        L.orc_adj_iterate.argtypes = [vp, C.c_int, dp]
        L.orc_adj_get_lambda.argtypes = [vp, dp]
        L.orc_adj_set_lambda.argtypes = [vp, dp]
        L.orc_adj_gradient.argtypes = [vp, dp]
        assert L.orc_config_size() == C.sizeof(_abi.Config), "ifl_config layout mismatch"
        assert L.orc_record_size() == C.sizeof(_abi.IterRecord), "ifl_iter_record layout mismatch"
        _lib = L
    return _lib


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


# Oracle-only field ids: the v and w copies of the momentum matrix (IFL_FIELD_MCOEF is u's).
FIELD_VCOEF = 100
FIELD_WCOEF = 101

# Planted mistakes (oracle/simple_oracle.hpp, FAULT_*): set only inside `planted_fault`.
FAULT_DEFECT_SIGN = 1
FAULT_DEFECT_NO_WEST = 2
FAULT_REFINE_SUBTRACT = 3
FAULT_INITFACE_RIGHT = 4
FAULT_V_MATRIX_ROW = 5
FAULT_MG_PROLONG_SIGN = 6
FAULT_MG_COARSE_EDGE = 7


class planted_fault:
    """Context manager: run the oracle with one planted mistake, then restore it."""

    def __init__(self, code):
        self.code = code

    def __enter__(self):
        lib().orc_set_fault(self.code)

    def __exit__(self, *exc):
        lib().orc_set_fault(0)
        return False


class OracleError(RuntimeError):
    pass


class Oracle:
    """One oracle case (FP32 by default, FP64 with fp64=True)."""

    def __init__(self, cfg, fp64=False):
        self.cfg = cfg.copy()
        self.fp64 = fp64
        self.h = C.c_void_p()
        self._check(lib().orc_create(C.byref(self.cfg), 1 if fp64 else 0, C.byref(self.h)))

    def _check(self, st):
        if st != 0:
            raise OracleError(f"oracle status {st}: {lib().orc_last_error().decode()}")
        return st

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self, field, shaped=True):
        n = lib().orc_field_size(self.h, field)
        out = np.empty(n, dtype=np.float64)
        self._check(lib().orc_get_field(self.h, field, out.ctypes.data_as(C.POINTER(C.c_double)), n))
        if not self.fp64:
            out = out.astype(np.float32)
        shape = self.cfg.shape(_abi.FIELD_MCOEF if field in (FIELD_VCOEF, FIELD_WCOEF) else field)
        return out.reshape(shape) if shaped else out

    def set(self, field, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel())
        self._check(lib().orc_set_field(self.h, field, a.ctypes.data_as(C.POINTER(C.c_double)), a.size))

    def set_face_profile(self, face, vals):
        if vals is None:
            self._check(lib().orc_set_face_profile(self.h, face, None, 0))
            return
        a = np.ascontiguousarray(np.asarray(vals, dtype=np.float64).ravel())
        self._check(lib().orc_set_face_profile(self.h, face, a.ctypes.data_as(C.POINTER(C.c_double)), a.size))

    def stage(self, stage):
        out = np.zeros(8, dtype=np.float64)
        self._check(lib().orc_run_stage(self.h, stage, out.ctypes.data_as(C.POINTER(C.c_double)), 8))
        return out

    def solve(self, max_iter, history=True):
        cap = max_iter if history else 0
        hist = (_abi.IterRecord * max(cap, 1))()
        n = C.c_int()
        conv = C.c_int()
        st = lib().orc_solve(self.h, max_iter, hist, cap, C.byref(n), C.byref(conv))
        recs = [hist[i] for i in range(min(n.value, cap))]
        return {"iterations": n.value, "converged": bool(conv.value), "status": st, "history": recs}

    def set_iter(self, it):
        lib().orc_set_iter(self.h, int(it))

    # ---- p' multigrid levels (IFL_PSOLVE_MG) ----
    MG_X, MG_B, MG_C, MG_E, MG_N, MG_T, MG_DG = range(7)

    def mg_build(self):
        """Levels of the current p' system; returns [(nx, ny, nz), ...] from level 0."""
        n = C.c_int()
        dims = (C.c_int * 48)()
        self._check(lib().orc_mg_build(self.h, C.byref(n), dims))
        self.mg_dims = [(dims[3 * l], dims[3 * l + 1], dims[3 * l + 2]) for l in range(n.value)]
        return self.mg_dims

    def mg_get(self, level, which):
        nx, ny, nz = self.mg_dims[level]
        out = np.empty(nx * ny * nz)
        self._check(lib().orc_mg_get(self.h, level, which, self._dp(out), out.size))
        return out.reshape(nz, ny, nx)

    def mg_set(self, level, which, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel())
        self._check(lib().orc_mg_set(self.h, level, which, self._dp(a), a.size))

    def mg_cycles(self, n):
        r = C.c_double()
        self._check(lib().orc_mg_cycles(self.h, int(n), C.byref(r)))
        return r.value

    def mg_smooth(self, level, n):
        self._check(lib().orc_mg_smooth(self.h, level, int(n)))

    # ---- discrete adjoint (oracle/adjoint_oracle.hpp) ----
    # Flat layouts: state = [u, v, w, p (N each), uf, vf, wf]; parameters = [face profiles
    # E, W, N, S, T, B (cells on each face), mu].
    def adj_setup(self, objective=0):
        self._check(lib().orc_adj_setup(self.h, objective))
        nv, npar = C.c_size_t(), C.c_size_t()
        self._check(lib().orc_adj_sizes(self.h, C.byref(nv), C.byref(npar)))
        self.adj_nvec, self.adj_npar = nv.value, npar.value

    @staticmethod
    def _dp(a):
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def adj_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.adj_nvec
        self._check(lib().orc_adj_set_weights(self.h, self._dp(w)))

    def adj_objective(self):
        return lib().orc_adj_objective(self.h)

    def adj_tangent(self, dx, dth):
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        dth = np.ascontiguousarray(dth, dtype=np.float64)
        out = np.zeros(self.adj_nvec)
        self._check(lib().orc_adj_tangent(self.h, self._dp(dx), self._dp(dth), self._dp(out)))
        return out

    def adj_transpose(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.zeros(self.adj_nvec)
        tb = np.zeros(self.adj_npar)
        self._check(lib().orc_adj_transpose(self.h, self._dp(lam), self._dp(out), self._dp(tb)))
        return out, tb

    def adj_iterate(self, n):
        sc = C.c_double()
        self._check(lib().orc_adj_iterate(self.h, int(n), C.byref(sc)))
        return sc.value

    def adj_lambda(self):
        out = np.zeros(self.adj_nvec)
        self._check(lib().orc_adj_get_lambda(self.h, self._dp(out)))
        return out

    def adj_set_lambda(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        self._check(lib().orc_adj_set_lambda(self.h, self._dp(lam)))

    def adj_gradient(self):
        out = np.zeros(self.adj_npar)
        self._check(lib().orc_adj_gradient(self.h, self._dp(out)))
        return out
