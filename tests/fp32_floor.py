"""The scaled-residual floor of the FP32 cavity against the grid size, on the oracle (the checker;
this script is not collected by pytest).

  python tests/fp32_floor.py --out floor.json [--sizes 32 64 128 256]

For each n: the 2D lid-driven cavity (Re 100, alpha 0.7/0.3) solved from rest in FP32 and in FP64
with the multigrid p' solve (one V-cycle, coarse-correction factor 1.4; the converged state does not
depend on the p' solver), run well past the point where FP32 stops improving.  Records the smallest
scaled residual each precision reaches, the iteration at which FP32 first came within 10 % of its
minimum, and which equation holds the floor.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for _p in (ROOT, os.path.join(ROOT, "incompflow-cuda_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

from incompflow import _abi  # noqa: E402
from oracle.oracle import Oracle, get_threads  # noqa: E402


def run(n, fp64, iters):
    c = _abi.cavity_config(n=n, p_refine=0, tol_residual=1e-14, max_iter=iters)
    c.stop_rule = _abi.STOP_NONE
    c.p_solver, c.mg_cycles, c.mg_over = _abi.PSOLVE_MG, 1, 1.4
    o = Oracle(c, fp64=fp64)
    t0 = time.perf_counter()
    h = []
    done = 0
    while done < iters:
        r = o.solve(min(5000, iters - done), history=True)
        h += r["history"]
        done += r["iterations"]
        if r["status"] != 0:
            break
    best = min(h, key=lambda x: x.scaled)
    first = next(i for i, x in enumerate(h) if x.scaled <= 1.1 * best.scaled) + 1
    # the equation holding the floor at the best record: residual over its normaliser
    parts = {"continuity": best.res_cont, "u": best.res_u, "v": best.res_v}
    return {"min_scaled": best.scaled, "first_within_10pct": first, "iterations": done,
            "s": time.perf_counter() - t0, "last_scaled": h[-1].scaled,
            "residuals_at_min": parts}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=[32, 64, 128, 256])
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    out = {"what": "smallest scaled residual of the 2D cavity (Re 100) in FP32 and FP64 on the oracle",
           "command": "python tests/fp32_floor.py --out <this file> --sizes " + " ".join(map(str, args.sizes)),
           "threads": get_threads(), "cases": []}
    for n in args.sizes:
        iters = max(4000, 60 * n)
        r32 = run(n, False, iters)
        r64 = run(n, True, iters)
        out["cases"].append({"n": n, "iterations": iters, "fp32": r32, "fp64": r64})
        print(f"[floor] n={n}: FP32 min {r32['min_scaled']:.3e} (from iteration {r32['first_within_10pct']}), "
              f"FP64 min {r64['min_scaled']:.3e}", flush=True)
        with open(args.out, "w") as fh:
            json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
