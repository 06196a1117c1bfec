"""Adjoint gradient of the 1024x256 adjoint channel against central differences of converged FP64
oracle solves (the oracle is the checker; this script is not collected by pytest).

Two phases, so the long FP64 solves can run on any host and the device part on a GPU:

  python tests/c5_adjoint_fd.py fd  --out FD.json      # FP64 oracle, host cores only
  python tests/c5_adjoint_fd.py gpu --fd FD.json --out profiles/adjoint_fd_fp64_c5_1024x256.json

fd : J = mean p(i=0) - mean p(i=nx-1) from FP64 oracle solves (multigrid p' solve, which does not
     change the converged state) converged to a 1e-10 scaled residual, at mu (1 +- h) and at the inlet
     profile (1 +- h), h = 1e-3 relative.  Each perturbed solve runs its first five iterations from
     rest (so its residual normalisers are a cold solve's), then continues from the converged base
     state.
gpu: the device's FP32 forward solve to a 1e-5 scaled residual and its discrete adjoint (until dJ/dmu
     and the profile gradient change by less than 1e-6 over 500 iterations), compared with the
     differences: dJ/dmu and the directional derivative along the profile itself.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for _p in (ROOT, os.path.join(ROOT, "incompflow-cuda_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

from incompflow import _abi  # noqa: E402

STATE = (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF)


def log(msg):
    print(msg, flush=True)


def case(nx, ny):
    c = _abi.adjoint_channel_config(nx, ny)
    return c, np.array(_abi.adjoint_inlet_profile(c))


def pressure_drop(p, nx):
    return float(p[:, :, 0].mean() - p[:, :, nx - 1].mean())


def fd_phase(args):
    from oracle.oracle import Oracle, get_threads
    c, prof = case(args.nx, args.ny)
    c.p_solver, c.p_refine, c.mg_over = _abi.PSOLVE_MG, 0, 1.0
    c.stop_rule, c.tol_residual, c.num_outer_iter = _abi.STOP_SCALED, args.tol, 10 ** 7

    def run(mu, p, base=None, label=""):
        c2 = c.copy()
        c2.viscosity = mu
        o = Oracle(c2, fp64=True)
        o.set_face_profile(_abi.WEST, p)
        t0 = time.perf_counter()
        done = 0
        if base is not None:
            done += o.solve(5, history=False)["iterations"]
            for f, a in base.items():
                o.set(f, a)
        while True:
            r = o.solve(2000, history=True)
            done += r["iterations"]
            log(f"[fd] {label}: {done} iterations, scaled {r['history'][-1].scaled:.3e}, "
                f"{time.perf_counter() - t0:.0f} s")
            if r["converged"] or r["status"] != 0:
                break
        assert r["converged"], f"{label}: FP64 solve did not converge (status {r['status']})"
        J = pressure_drop(o.get(_abi.FIELD_P), c.nx)
        st = {f: o.get(f) for f in STATE}
        o.close()
        return J, done, time.perf_counter() - t0, st

    mu, h = c.viscosity, args.h
    J0, n0, t0, base = run(mu, prof, None, "base")
    out = {"grid": [c.nx, c.ny], "h": h, "tol": args.tol, "threads": get_threads(), "J": J0,
           "solves": {"base": {"iterations": n0, "s": t0}}}
    for name, m, p in (("mu+", mu * (1 + h), prof), ("mu-", mu * (1 - h), prof),
                       ("prof+", mu, prof * (1 + h)), ("prof-", mu, prof * (1 - h))):
        J, n, t, _ = run(m, p, base, name)
        out["solves"][name] = {"J": J, "iterations": n, "s": t}
    s = out["solves"]
    out["dJ_dmu_fd"] = (s["mu+"]["J"] - s["mu-"]["J"]) / (2 * h * mu)
    out["dir_profile_fd"] = (s["prof+"]["J"] - s["prof-"]["J"]) / (2 * h)
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    log(f"[fd] dJ/dmu {out['dJ_dmu_fd']:.8e}, profile direction {out['dir_profile_fd']:.8e}")


def gpu_phase(args):
    import incompflow
    with open(args.fd) as fh:
        fd = json.load(fh)
    c, prof = case(*fd["grid"])
    c.stop_rule, c.tol_residual, c.num_outer_iter = _abi.STOP_SCALED, args.forward_tol, 400000
    c.sweep_block = 4
    s = incompflow.Solver(c)
    s.set_face_profile(_abi.WEST, prof)
    t0 = time.perf_counter()
    rf = s.solve(0, history=False)
    tf = time.perf_counter() - t0
    log(f"[gpu] forward {rf['iterations']} iterations, {tf:.1f} s, converged {rf['converged']}")
    t0 = time.perf_counter()
    ra = s.adjoint(_abi.OBJ_PRESSURE_DROP, _abi.WEST, tol=1e-7, grad_tol=1e-6, check_every=500, max_iter=200000)
    ta = time.perf_counter() - t0
    J = pressure_drop(s.get(_abi.FIELD_P), c.nx)
    s.close()
    adj_dir = float(np.dot(prof, ra["dJ_dprofile"]))
    out = {
        "what": "FP32 device adjoint gradient of the 1024x256 adjoint channel against central differences "
                "of FP64 oracle solves converged to a 1e-10 scaled residual",
        "command": "python tests/c5_adjoint_fd.py fd --out FD.json (host); "
                   "python tests/c5_adjoint_fd.py gpu --fd FD.json --out <this file> (MI355X)",
        "grid": fd["grid"], "fd_step_relative": fd["h"], "fd_tol": fd["tol"], "fd_threads": fd["threads"],
        "fd_solves": fd["solves"],
        "forward": {"tol": args.forward_tol, "iterations": rf["iterations"], "converged": rf["converged"], "s": tf},
        "adjoint": {"iterations": ra["iterations"], "converged": ra["converged"], "s": ta,
                    "grad_change": ra["grad_change"]},
        "J_device_fp32": J, "J_oracle_fp64": fd["J"], "J_rel_err": abs(J - fd["J"]) / abs(fd["J"]),
        "dJ_dmu_adjoint": ra["dJ_dmu"], "dJ_dmu_fd": fd["dJ_dmu_fd"],
        "dJ_dmu_rel_err": abs(ra["dJ_dmu"] - fd["dJ_dmu_fd"]) / abs(fd["dJ_dmu_fd"]),
        "dir_profile_adjoint": adj_dir, "dir_profile_fd": fd["dir_profile_fd"],
        "dir_profile_rel_err": abs(adj_dir - fd["dir_profile_fd"]) / abs(fd["dir_profile_fd"]),
        "adjoint_over_forward": ta / tf,
    }
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    log(f"[gpu] dJ/dmu rel err {out['dJ_dmu_rel_err']:.2e}, profile direction rel err {out['dir_profile_rel_err']:.2e}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("phase", choices=["fd", "gpu"])
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--h", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--forward-tol", type=float, default=1e-5)
    ap.add_argument("--fd", default=None)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    fd_phase(args) if args.phase == "fd" else gpu_phase(args)


if __name__ == "__main__":
    main()
