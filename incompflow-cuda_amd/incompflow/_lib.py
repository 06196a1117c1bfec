"""ctypes binding of libincompflow.so (include/incompflow.h).

The library is the product: it must be built (``make -C incompflow-cuda_amd/csrc``) and a GPU
must be present; there is no CPU fallback, and every failing call raises IflError with the
library's own message.
"""
import ctypes as C
import os

import numpy as np

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
# IFL_LIBRARY selects another build of the same sources (the host-sanitizer build, csrc/_san/)
LIB_PATH = os.environ.get("IFL_LIBRARY") or os.path.join(_HERE, "libincompflow.so")

# (name, restype, argtypes) for every entry point declared in include/incompflow.h
EXPORTS = [
    ("ifl_last_error", C.c_char_p, []),
    ("ifl_abi_version", C.c_int, []),
    ("ifl_config_default", None, [C.POINTER(_abi.Config)]),
    ("ifl_create", C.c_int, [C.POINTER(_abi.Config), C.POINTER(C.c_void_p)]),
    ("ifl_destroy", C.c_int, [C.c_void_p]),
    ("ifl_set_face_profile", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_size_t]),
    ("ifl_solve", C.c_int, [C.c_void_p, C.c_int, C.POINTER(_abi.Report), C.POINTER(_abi.IterRecord), C.c_size_t]),
    ("ifl_get_field", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_size_t]),
    ("ifl_set_field", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_size_t]),
    ("ifl_field_size", C.c_size_t, [C.c_void_p, C.c_int]),
    ("ifl_write_vtk", C.c_int, [C.c_void_p, C.c_char_p]),
    ("ifl_save_state", C.c_int, [C.c_void_p, C.c_char_p]),
    ("ifl_load_state", C.c_int, [C.c_void_p, C.c_char_p]),
    ("ifl_run_stage", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_size_t]),
    ("ifl_time_kernel", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    ("ifl_autotune", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(_abi.TuneResult)]),
    ("ifl_adjoint_opts_default", None, [C.POINTER(_abi.AdjointOpts)]),
    ("ifl_adjoint", C.c_int, [C.c_void_p, C.POINTER(_abi.AdjointOpts), C.c_int, C.POINTER(C.c_double), C.c_size_t,
                              C.POINTER(C.c_double), C.POINTER(_abi.AdjointReport)]),
    ("ifl_adjoint_lambda", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_size_t]),
    ("ifl_batch_create", C.c_int, [C.POINTER(_abi.Config), C.c_int, C.POINTER(C.c_void_p)]),
    ("ifl_batch_destroy", C.c_int, [C.c_void_p]),
    ("ifl_batch_size", C.c_int, [C.c_void_p]),
    ("ifl_batch_case", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    ("ifl_batch_solve", C.c_int, [C.c_void_p, C.c_int, C.POINTER(_abi.Report)]),
    ("ifl_batch_autotune", C.c_int, [C.c_void_p, C.c_int, C.POINTER(_abi.TuneResult)]),
    ("ifl_solve_noalloc", C.c_int, [C.POINTER(_abi.Config), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_int,
                                    C.POINTER(_abi.Report), C.POINTER(_abi.IterRecord), C.c_size_t]),
]

_lib = None


class IflError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"incompflow status {status}: {msg}")
        self.status = status


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IflError(-1, f"{LIB_PATH} is missing: build it with `make -C incompflow-cuda_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in EXPORTS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.ifl_abi_version() != _abi.IFL_ABI_VERSION:
            raise IflError(-1, "ABI version mismatch")
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise IflError(st, lib().ifl_last_error().decode())
    return st


class Solver:
    """One case on the GPU: the reference's Solver (include/Solver.h:46-52) over the C-ABI."""

    def __init__(self, cfg, _handle=None, _owner=None):
        self.cfg = cfg.copy()
        self._owner = _owner          # a Batch that owns the handle, or None
        if _handle is not None:
            self.h = _handle
        else:
            self.h = C.c_void_p()
            _check(lib().ifl_create(C.byref(self.cfg), C.byref(self.h)))

    def close(self):
        if self.h and self._owner is None:
            lib().ifl_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def solve(self, max_iter=0, history=True):
        """Run outer iterations; returns dict(iterations, converged, status, elapsed_ms, history)."""
        n = max_iter if max_iter > 0 else self.cfg.num_outer_iter
        cap = n if history else 0
        hist = (_abi.IterRecord * max(cap, 1))()
        rep = _abi.Report()
        st = lib().ifl_solve(self.h, max_iter, C.byref(rep), hist, cap)
        if st not in (0, _abi.IFL_ERR_DIVERGED, _abi.IFL_ERR_NOT_CONVERGED):
            _check(st)
        recs = [hist[i] for i in range(min(rep.iterations, cap))]
        return {"iterations": rep.iterations, "converged": bool(rep.converged), "status": st,
                "elapsed_ms": rep.elapsed_ms, "final_scaled": rep.final_scaled,
                "sweeps_total": rep.sweeps_total, "history": recs}

    def get(self, field, shaped=True):
        n = self.cfg.field_size(field)
        out = np.empty(n, dtype=np.float32)
        _check(lib().ifl_get_field(self.h, field, out.ctypes.data_as(C.POINTER(C.c_float)), n))
        return out.reshape(self.cfg.shape(field)) if shaped else out

    def set(self, field, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float32).ravel())
        _check(lib().ifl_set_field(self.h, field, a.ctypes.data_as(C.POINTER(C.c_float)), a.size))

    def set_face_profile(self, face, vals):
        if vals is None:
            _check(lib().ifl_set_face_profile(self.h, face, None, 0))
            return
        a = np.ascontiguousarray(np.asarray(vals, dtype=np.float64).ravel())
        _check(lib().ifl_set_face_profile(self.h, face, a.ctypes.data_as(C.POINTER(C.c_double)), a.size))

    def stage(self, stage):
        out = np.zeros(8, dtype=np.float64)
        _check(lib().ifl_run_stage(self.h, stage, out.ctypes.data_as(C.POINTER(C.c_double)), 8))
        return out

    def time_kernel(self, kernel, reps=50):
        """Mean device time per launch (microseconds) and launches per repetition."""
        us = C.c_double()
        per = C.c_int()
        _check(lib().ifl_time_kernel(self.h, kernel, reps, C.byref(us), C.byref(per)))
        return us.value, per.value

    def write_vtk(self, path):
        _check(lib().ifl_write_vtk(self.h, os.fsencode(path)))

    def save_state(self, path):
        """Binary state file: fields and outer-iteration monitor (ifl_save_state)."""
        _check(lib().ifl_save_state(self.h, os.fsencode(path)))

    def load_state(self, path):
        """Continue from (or warm start with) a state file of the same grid (ifl_load_state)."""
        _check(lib().ifl_load_state(self.h, os.fsencode(path)))

    def adjoint(self, objective=_abi.OBJ_PRESSURE_DROP, face=_abi.WEST, weights=None, max_iter=0, tol=0.0,
                warm_start=False, check_every=0, profile_gradient=True, grad_tol=0.0, tile=0):
        """Gradient of the objective at the current (converged) state w.r.t. the normal-velocity
        profile of `face` (an inlet) and mu (ifl_adjoint).  weights: dict field id -> array
        (OBJ_LINEAR).  tol < 0 runs exactly max_iter adjoint iterations.  tile: 2D transposed-sweep
        tile (0 = by grid size, 1..6 as ifl_adjoint_opts.tile)."""
        o = _abi.AdjointOpts()
        lib().ifl_adjoint_opts_default(C.byref(o))
        o.objective = objective
        if max_iter:
            o.max_iter = max_iter
        if tol:
            o.tol = tol
        o.grad_tol = grad_tol
        o.warm_start = 1 if warm_start else 0
        if check_every:
            o.check_every = check_every
        o.tile = tile
        keep = []
        if weights:
            for q, fid in enumerate((_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W, _abi.FIELD_P)):
                if fid in weights:
                    a = np.ascontiguousarray(np.asarray(weights[fid], dtype=np.float32).ravel())
                    keep.append(a)
                    o.weights[q] = a.ctypes.data_as(C.POINTER(C.c_float))
        c = self.cfg
        nprof = 0
        gp = None
        if profile_gradient:
            nprof = c.ny * c.nz if face < 2 else c.nx * c.nz if face < 4 else c.nx * c.ny
            gp = np.zeros(nprof, dtype=np.float64)
        gmu = C.c_double()
        rep = _abi.AdjointReport()
        _check(lib().ifl_adjoint(self.h, C.byref(o), face,
                                 gp.ctypes.data_as(C.POINTER(C.c_double)) if gp is not None else None,
                                 nprof, C.byref(gmu), C.byref(rep)))
        return {"dJ_dprofile": gp, "dJ_dmu": gmu.value, "iterations": rep.iterations,
                "converged": bool(rep.converged), "J": rep.J, "final_scaled": rep.final_scaled,
                "elapsed_ms": rep.elapsed_ms, "grad_change": rep.grad_change}

    def adjoint_lambda(self, field, shaped=True):
        n = self.cfg.field_size(field)
        out = np.empty(n, dtype=np.float32)
        _check(lib().ifl_adjoint_lambda(self.h, field, out.ctypes.data_as(C.POINTER(C.c_float)), n))
        return out.reshape(self.cfg.shape(field)) if shaped else out

    def autotune(self, what=_abi.TUNE_LAUNCH | _abi.TUNE_SWEEPS, trial_iters=0):
        """Pick tile, sweep_block(_mom) and sweeps per outer iteration by measurement (ifl_autotune);
        the choice is applied to this solver and mirrored into self.cfg."""
        r = _abi.TuneResult()
        _check(lib().ifl_autotune(self.h, what, trial_iters, C.byref(r)))
        if self.cfg.dim == 2:
            self.cfg.tile = r.tile
        self.cfg.sweep_block = r.sweep_block
        self.cfg.sweep_block_mom = r.sweep_block_mom
        for q in range(4):
            self.cfg.n_iter[q] = r.n_iter[q]
        self.cfg.p_solver = r.p_solver
        self.cfg.mg_cycles = r.mg_cycles
        self.cfg.mg_over = r.mg_over
        if r.p_solver == _abi.PSOLVE_MG:
            self.cfg.p_refine = 0
        return {"tile": r.tile, "sweep_block": r.sweep_block, "sweep_block_mom": r.sweep_block_mom,
                "n_iter": list(r.n_iter), "p_solver": r.p_solver, "mg_cycles": r.mg_cycles, "mg_over": r.mg_over,
                "us_per_iter": r.us_per_iter, "decades_per_s": r.decades_per_s, "trials": r.trials}


class Batch:
    """Many independent cases solved together (ifl_batch_* in include/incompflow.h).

    Cases may differ in grid size, domain, boundary conditions, fluid properties, relaxation,
    tolerances and stop rule; dim, smoother, n_iter, p_refine, sweep_block and tile must agree.
    ``cases[i]`` is a Solver view of case i (fields, profiles, VTK) owned by the batch."""

    def __init__(self, cfgs):
        cfgs = [c.copy() for c in cfgs]
        arr = (_abi.Config * len(cfgs))(*cfgs)
        self.h = C.c_void_p()
        _check(lib().ifl_batch_create(arr, len(cfgs), C.byref(self.h)))
        self.cases = []
        for i, c in enumerate(cfgs):
            hc = C.c_void_p()
            _check(lib().ifl_batch_case(self.h, i, C.byref(hc)))
            self.cases.append(Solver(c, _handle=hc, _owner=self))

    def __len__(self):
        return len(self.cases)

    def close(self):
        if self.h:
            for s in self.cases:
                s.h = C.c_void_p()
            lib().ifl_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def solve(self, max_iter=0):
        """Run until every case stopped; returns (status, [dict per case])."""
        reps = (_abi.Report * len(self.cases))()
        st = lib().ifl_batch_solve(self.h, max_iter, reps)
        if st not in (0, _abi.IFL_ERR_DIVERGED, _abi.IFL_ERR_NOT_CONVERGED):
            _check(st)
        out = [{"iterations": r.iterations, "converged": bool(r.converged), "status": r.status,
                "elapsed_ms": r.elapsed_ms, "final_scaled": r.final_scaled, "sweeps_total": r.sweeps_total}
               for r in reps]
        return st, out

    def autotune(self, trial_iters=0):
        """Pick the tile, sweep_block and sweep_block_mom shared by every case by measurement
        (ifl_batch_autotune); the choice is applied to the batch and mirrored into each case's cfg."""
        r = _abi.TuneResult()
        _check(lib().ifl_batch_autotune(self.h, trial_iters, C.byref(r)))
        for s in self.cases:
            s.cfg.tile, s.cfg.sweep_block, s.cfg.sweep_block_mom = r.tile, r.sweep_block, r.sweep_block_mom
        return {"tile": r.tile, "sweep_block": r.sweep_block, "sweep_block_mom": r.sweep_block_mom,
                "us_per_iter": r.us_per_iter, "trials": r.trials}


def solve_noalloc(cfg, u, v, p, w=None, temp=None, face_profiles=None, max_iter=0, stream=None, history=False):
    """One solve with no device allocation (ifl_solve_noalloc) on caller-owned torch CUDA tensors.

    u, v, p (w in 3D, temp with energy on) are contiguous float32 tensors of nx*ny*nz elements on
    one device, read as the initial state and overwritten with the result.  face_profiles maps a
    face to a float32 CUDA tensor of that face's cell count.  stream: a torch.cuda.Stream (default:
    torch's current stream on the fields' device).  Returns the dict of Solver.solve."""
    import torch
    n = cfg.nx * cfg.ny * cfg.nz
    dev = u.device if isinstance(u, torch.Tensor) else None

    def ptr(t, name, count):
        if t is None:
            return None
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                and t.numel() == count and t.device == dev):
            raise IflError(_abi.IFL_ERR_ARG, f"{name} must be a contiguous float32 CUDA tensor of {count} "
                                             f"elements on the device of u")
        return C.c_void_p(t.data_ptr())

    args = [ptr(t, nm, n) for t, nm in ((u, "u"), (v, "v"), (w, "w"), (p, "p"), (temp, "temp"))]
    prof = None
    if face_profiles:
        prof = (C.c_void_p * 6)()
        for face, t in face_profiles.items():
            cells = (cfg.ny * cfg.nz if face < 2 else cfg.nx * cfg.nz if face < 4 else cfg.nx * cfg.ny)
            prof[face] = ptr(t, f"face profile {face}", cells)
    cap = (max_iter if max_iter > 0 else cfg.num_outer_iter) if history else 0
    hist = (_abi.IterRecord * max(cap, 1))()
    rep = _abi.Report()
    with torch.cuda.device(dev):
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        st = lib().ifl_solve_noalloc(C.byref(cfg), C.c_void_p(s.cuda_stream), args[0], args[1], args[2], args[3],
                                     args[4], prof, max_iter, C.byref(rep), hist, cap)
    if st not in (0, _abi.IFL_ERR_DIVERGED, _abi.IFL_ERR_NOT_CONVERGED):
        _check(st)
    recs = [hist[i] for i in range(min(rep.iterations, cap))]
    return {"iterations": rep.iterations, "converged": bool(rep.converged), "status": st,
            "elapsed_ms": rep.elapsed_ms, "final_scaled": rep.final_scaled,
            "sweeps_total": rep.sweeps_total, "history": recs}
