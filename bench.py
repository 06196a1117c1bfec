"""Benchmark: steady SIMPLE loop, 2D lid-driven cavity 1024x1024 (Re=100), FP32, on one GPU.

One step = one SIMPLE outer iteration (momentum assembly, u/v smoother sweeps, Rhie-Chow,
p' assembly, p' sweeps + one FP64-defect refinement cycle, correction, FP64 norms and stop test)
over the whole grid, run through the library's captured hipGraph.  Prints one JSON line.

value      : smoother cell-updates per second over the timed steps (cells x sweeps executed / s,
             the whole outer iteration's time in the denominator)
roofline   : the dominant kernel, the fused red-black p' sweep (k_rb_sweep2d, 5 sweeps per launch),
             timed with HIP events on the library's stream.  achieved = the launch's compulsory
             bytes (the field read and written once, b' and the +x/+y coefficient planes: 20 B per
             cell for all its sweeps) / launch time; traffic = the HBM bytes the rocprofv3 counters
             measured for the same launch (profiles/); the per-sweep streaming byte model (32 B per
             cell and sweep) is reported beside it, not as frac
time_to_residual_s: a separate solve from rest to the scaled residual --target (1e-4).  A 1e-5 scaled
             residual is below the FP32 floor of this grid (time_to_1e-5: reached false, with the
             floor measured by continuing the multigrid solve past 1e-4)
refine_ab  : the same solve from rest without the FP64-defect refinement cycle (p_refine 0): its
             iterations and time to the target and its time per outer iteration (unless --no-refine-ab)
p_solver_mg: the same cavity with the multigrid p' solve, its strength chosen by the autotuner
             (IFL_TUNE_PSOLVE), solved from rest to the same residual
fp64_error : a 256x256 cavity (and the 256x64 adjoint channel) converged on the GPU in FP32 against
             the FP64 oracle converged to 1e-10: max |x - x64| / max |x64| for u, v and p (mean-free),
             as measured by tests/test_gpu_fp64_error.py and summarised under profiles/
cpu_baseline: the FP32 CPU oracle (OpenMP, host cores) on a bounded number of outer iterations, with
             the host's CPU model and core counts and a 1-thread figure
batch      : the seeded 64-case uneven batch (2D cavities and channels, 64..512 cells a side,
             Re 10..1000, SplitMix64 seed 20261203) solved to the same scaled residual in one batch
             with launch settings chosen by a batch-level trial, against the same cases solved one
             after another in full (unless --no-batch).  Each case of the batch equals its single
             solve bit for bit, so both runs execute the same iterations.
energy     : the same cavity with heated walls and the temperature equation on: ms per outer
             iteration with and without it, and the temperature sweep launch (unless --no-energy).
channel_3000: the reference's default 512x128 channel, 3000 outer iterations on the GPU and on the FP32
             oracle (host cores), timed directly.
adjoint    : the 1024x256 channel with a noisy parabolic inlet solved to the scaled residual
             --adjoint-forward-target, then the discrete adjoint of the pressure drop until dJ/dmu and
             the profile gradient settle; wall times of both, and central differences of FP32 device
             solves (warm-started from the converged state file) along mu and along the profile.
The GPU and the FP32 oracle produce the same iterates bit for bit, so the CPU time to the same
residual is the oracle's time per iteration times the same iteration count.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "incompflow-cuda_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
METRIC = "time to 1e-5 residual & cell-updates/s (%HBM peak), plain vs batched+adjoint"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# the CPU oracle's OpenMP threads, one per physical core, kept there (read when the OpenMP runtime
# starts, so before torch loads it)
os.environ.setdefault("OMP_PLACES", "cores")
os.environ.setdefault("OMP_PROC_BIND", "close")

try:   # imported before any device work: a first import between the warmup and the timed steps
    import torch  # leaves the GPU idle for seconds, and the timed steps then start at idle clocks
except ImportError:
    torch = None


def device_sync():
    if torch is not None and torch.cuda.is_available():
        torch.cuda.synchronize()


def sweep_bytes_per_cell(dim):
    # compulsory FP32 traffic of one smoother sweep: (1+2dim) coefficients + source + phi read/write
    return 4 * ((1 + 2 * dim) + 1 + 2)


def batch_leg(args, incompflow, _abi):
    cfgs = _abi.uneven_batch_configs(tol_residual=args.batch_target, max_iter=args.batch_max_iter)
    for c in cfgs:
        c.sweep_block = args.sweep_block_batch
    cells = sum(c.nx * c.ny for c in cfgs)
    log(f"[bench] batch: {len(cfgs)} cases, {cells} cells, to scaled residual {args.batch_target:g}")
    b = incompflow.Batch(cfgs)
    tune = b.autotune(args.batch_trial_iters)
    log(f"[bench] batch launch settings by the batch autotuner ({tune['trials']} trials): tile {tune['tile']}, "
        f"sweep_block {tune['sweep_block']}, sweep_block_mom {tune['sweep_block_mom']}, {tune['us_per_iter']:.1f} us/iteration")
    for c in cfgs:
        c.tile, c.sweep_block, c.sweep_block_mom = tune["tile"], tune["sweep_block"], tune["sweep_block_mom"]
    device_sync()
    t0 = time.perf_counter()
    st, reps = b.solve()
    device_sync()
    tb = time.perf_counter() - t0
    b.close()
    its = [r["iterations"] for r in reps]
    conv = sum(1 for r in reps if r["converged"])
    log(f"[bench] batch {tb:.2f} s, {conv}/{len(cfgs)} converged, max {max(its)} iterations")
    seq = 0.0
    if args.batch_sequential == "none":
        kind = "none: not run"
    elif args.batch_sequential == "full":
        for i, c in enumerate(cfgs):
            s = incompflow.Solver(c)
            t0 = time.perf_counter()
            r = s.solve(0, history=False)
            device_sync()
            seq += time.perf_counter() - t0
            s.close()
            if r["iterations"] != its[i] or r["converged"] != reps[i]["converged"]:
                raise RuntimeError(f"case {i}: single solve ran {r['iterations']} iterations, batch {its[i]}")
            if i % 8 == 7:
                log(f"[bench]   sequential: {i + 1} cases, {seq:.1f} s")
        kind = "measured: every case solved alone, with the same settings, to the same residual"
    else:
        for i, c in enumerate(cfgs):
            c2 = c.copy()
            c2.stop_rule = _abi.STOP_NONE
            s = incompflow.Solver(c2)
            s.solve(20, history=False)
            device_sync()
            t0 = time.perf_counter()
            s.solve(args.batch_probe_iters, history=False)
            device_sync()
            seq += (time.perf_counter() - t0) / args.batch_probe_iters * its[i]
            s.close()
        kind = (f"estimated: each case's single-solve wall time per iteration "
                f"({args.batch_probe_iters} iterations) x its iteration count")
    if args.batch_sequential != "none":
        log(f"[bench] sequential {seq:.2f} s ({kind.split(':')[0]}), speedup {seq / tb:.2f}x")
    stalled = [{"case": i, "nx": cfgs[i].nx, "ny": cfgs[i].ny, "re": round(1.0 / cfgs[i].viscosity, 1),
                "kind": "cavity" if cfgs[i].bc_type[_abi.WEST] == _abi.WALL else "channel",
                "final_scaled": reps[i]["final_scaled"]} for i in range(len(cfgs)) if not reps[i]["converged"]]
    return {"cases": len(cfgs), "cells": cells, "residual_target": args.batch_target,
            "max_iter": args.batch_max_iter, "converged": conv, "not_converged": stalled,
            "iterations_max": max(its), "iterations_sum": sum(its), "batch_s": tb, "sequential_s": seq,
            "sequential_kind": kind, "speedup": seq / tb if seq else None, "launch_settings": tune,
            "cell_updates_per_s": sum(r["sweeps_total"] for r in reps) / tb}


# algorithmic bytes per cell of one launch (FP32 words x 4): the compulsory reads and writes
KERNEL_WORDS_3D = {"mom_assemble": 14, "rhie_chow": 8, "pcorr_assemble": 13, "update": 16}


def cavity3d_leg(args, incompflow, _abi):
    """3D lid-driven cavity 256^3 (Re 400): time per outer iteration and per kernel, with each
    kernel's algorithmic GB/s; the FP32 oracle's time per iteration on the host cores beside it."""
    n = args.c3_n
    c = _abi.cavity3d_config(n=n)
    c.p_refine = 0
    c.stop_rule = _abi.STOP_NONE
    cells = n ** 3
    log(f"[bench] cavity3d {n}^3: {args.c3_steps} outer iterations")
    s = incompflow.Solver(c)
    s.solve(3, history=False)
    device_sync()
    t0 = time.perf_counter()
    r = s.solve(args.c3_steps, history=True)
    device_sync()
    wall = time.perf_counter() - t0
    sweeps = sum(sum(rec.sweeps) for rec in r["history"]) / args.c3_steps
    ms = 1000.0 * wall / args.c3_steps
    kern = {}
    reps = 20
    us, _ = s.time_kernel(_abi.KERNEL_SWEEP, reps=reps)
    # p' sweep: field in and out, b', the +x/+y/+z coefficient planes (24 B/cell); the survey's
    # byte model of a stored 7-point sweep is 40 B/cell
    kern["pcorr_sweep"] = {"us": us, "own_gbs": 24.0 * cells / (us * 1e3), "model_gbs": 40.0 * cells / (us * 1e3)}
    us, _ = s.time_kernel(_abi.KERNEL_MOM_SWEEP, reps=reps)
    # u, v, w together: 3 fields in and out, 3 sources, the 7 coefficient planes (64 B/cell)
    kern["momentum_sweep_uvw"] = {"us": us, "own_gbs": 64.0 * cells / (us * 1e3),
                                  "model_gbs": 3 * 40.0 * cells / (us * 1e3)}
    for name, kid in (("mom_assemble", _abi.KERNEL_MOM_ASSEMBLE), ("rhie_chow", _abi.KERNEL_RHIE_CHOW),
                      ("pcorr_assemble", _abi.KERNEL_PCORR_ASSEMBLE), ("update", _abi.KERNEL_UPDATE)):
        us, _ = s.time_kernel(kid, reps=reps)
        kern[name] = {"us": us, "model_gbs": 4.0 * KERNEL_WORDS_3D[name] * cells / (us * 1e3)}
    us, _ = s.time_kernel(_abi.KERNEL_PSOLVE, reps=5)
    kern["p_solve_sweeps"] = {"us": us}
    s.close()
    log(f"[bench] cavity3d {ms:.3f} ms/outer iteration ({sweeps:.0f} sweeps), p' sweep "
        f"{kern['pcorr_sweep']['us']:.1f} us, u+v+w sweep {kern['momentum_sweep_uvw']['us']:.1f} us")
    # the same case with the multigrid p' solve (one V-cycle, coarse-correction factor 1.4)
    cm = c.copy()
    cm.p_solver, cm.mg_cycles, cm.mg_over = _abi.PSOLVE_MG, 1, 1.4
    s = incompflow.Solver(cm)
    s.solve(3, history=False)
    device_sync()
    t0 = time.perf_counter()
    s.solve(args.c3_steps, history=False)
    device_sync()
    ms_mg = 1000.0 * (time.perf_counter() - t0) / args.c3_steps
    us_mg, _ = s.time_kernel(_abi.KERNEL_PSOLVE, reps=5)
    s.close()
    log(f"[bench] cavity3d multigrid p' solve: {ms_mg:.3f} ms/outer iteration, V-cycle {us_mg:.1f} us")
    out = {"workload": f"3D lid-driven cavity {n}^3, Re=400, FP32 RB-GS 10/10/10/20 sweeps, alpha 0.7/0.3, "
                       "FP64 reductions", "cells": cells, "ms_per_step": ms, "sweeps_per_step": sweeps,
           "cell_updates_per_s": cells * sweeps / (ms * 1e-3), "kernels": kern, "cpu_ms_per_step": None,
           "multigrid": {"ms_per_step": ms_mg, "p_solve_us": us_mg, "mg_cycles": 1, "mg_over": 1.4}}
    # time to residual: measured by the opt-in --c3-converge run and kept under profiles/
    cfile = os.path.join(ROOT, "profiles", f"cavity3d_{n}_time_to_residual.json")
    if args.c3_converge:
        out["time_to_residual"] = cavity3d_converge(args, incompflow, _abi, n)
        if args.c3_converge_out:
            with open(args.c3_converge_out, "w") as fh:
                json.dump(out["time_to_residual"], fh, indent=1)
    else:
        for key, fn in (("time_to_residual", cfile), ("time_to_residual_mg", cfile.replace(".json", "_mg.json"))):
            if os.path.exists(fn):
                with open(fn) as fh:
                    out[key] = json.load(fh)
                out[key]["source"] = os.path.relpath(fn, ROOT)
    if not args.no_cpu and args.c3_cpu_iters > 0:
        from oracle import oracle as orc
        o = orc.Oracle(c)
        o.solve(1, history=False)
        t0 = time.perf_counter()
        o.solve(args.c3_cpu_iters, history=False)
        out["cpu_ms_per_step"] = 1000.0 * (time.perf_counter() - t0) / args.c3_cpu_iters
        out["cpu_threads"] = orc.get_threads()
        out["gpu_vs_cpu_per_step"] = out["cpu_ms_per_step"] / ms
        o.close()
        log(f"[bench] cavity3d cpu oracle {out['cpu_ms_per_step']:.1f} ms/iteration")
    return out


def channel_leg(args, incompflow, _abi):
    """The reference's default case (512x128 channel, west inlet u = 0.001, east outlet, alpha 0.7/0.3,
    10/10/20 sweeps): exactly 3000 outer iterations from rest, the GPU against the FP32 oracle on the
    host cores doing the same 3000 (the two produce the same fields bit for bit: tests/test_gpu_parity.py)."""
    c = _abi.channel_config(512, 128, 3000)
    s = incompflow.Solver(c)
    tune = None
    if not args.no_tune:
        # launch settings by the autotuner (a small grid: sweeps per launch set the launch count);
        # the oracle runs the same settings, so both keep producing the same fields
        tune = s.autotune(_abi.TUNE_LAUNCH)
        c.tile, c.sweep_block, c.sweep_block_mom = s.cfg.tile, s.cfg.sweep_block, s.cfg.sweep_block_mom
        s.close()
        s = incompflow.Solver(c)
    device_sync()
    t0 = time.perf_counter()
    r = s.solve(3000, history=False)
    device_sync()
    tg = time.perf_counter() - t0
    s.close()
    out = {"workload": "2D channel 512x128 (the reference's config.h), 3000 outer iterations, RB-GS 10/10/20, "
                       "alpha 0.7/0.3", "iterations": r["iterations"], "gpu_s": tg, "gpu_ms_per_step": 1000 * tg / 3000,
           "gpu_device_ms_per_step": r["elapsed_ms"] / 3000, "cpu_s": None,
           "launch_settings": {"tile": c.tile, "sweep_block": c.sweep_block, "sweep_block_mom": c.sweep_block_mom,
                               "from": "the autotuner (IFL_TUNE_LAUNCH)" if tune else "defaults"}}
    if not args.no_cpu:
        from oracle import oracle as orc
        o = orc.Oracle(c)
        t0 = time.perf_counter()
        ro = o.solve(3000, history=False)
        tc = time.perf_counter() - t0
        o.close()
        out.update({"cpu_s": tc, "cpu_iterations": ro["iterations"], "cpu_threads": orc.get_threads(),
                    "gpu_vs_cpu": tc / tg})
    log(f"[bench] channel 512x128, 3000 iterations: GPU {tg:.2f} s"
        + (f", CPU oracle {out['cpu_s']:.1f} s ({out['gpu_vs_cpu']:.0f}x)" if out["cpu_s"] else ""))
    return out


def cavity3d_converge(args, incompflow, _abi, n):
    """The 3D cavity from rest towards the scaled residuals 1e-3, 1e-4 and 1e-5 (stop rule off, so
    the run continues to its budget and shows where FP32 stalls): outer iterations and device time
    at which each was first reached, and the smallest scaled residual seen."""
    c = _abi.cavity3d_config(n=n, tol_residual=1e-12, max_iter=10 ** 7)
    c.p_refine = 0
    c.stop_rule = _abi.STOP_NONE
    c.num_outer_iter = 10 ** 7
    chosen = None
    if args.c3_mg_over < 0:
        # the p' solver by the autotuner's time to 1e-4 on a 4x coarser copy of the cavity
        ct = c.copy()
        ct.stop_rule, ct.tol_residual = _abi.STOP_SCALED, 1e-4
        t = incompflow.Solver(ct)
        chosen = t.autotune(_abi.TUNE_PSOLVE_TARGET)
        t.close()
        c.p_solver, c.mg_cycles, c.mg_over = chosen["p_solver"], max(1, chosen["mg_cycles"]), chosen["mg_over"]
        log(f"[bench] cavity3d p' solver by the autotuner: {chosen['p_solver']} (mg_over {chosen['mg_over']})")
    elif args.c3_mg_over > 0:
        c.p_solver, c.mg_cycles, c.mg_over = _abi.PSOLVE_MG, 1, args.c3_mg_over
    targets = [1e-3, 1e-4, 1e-5]
    log(f"[bench] cavity3d {n}^3 from rest towards {targets} (budget {args.c3_converge_budget_s:.0f} s)")
    s = incompflow.Solver(c)
    hit, best, done, dev_s = solve_tracking(s, targets, 2000, args.c3_converge_budget_s, extra_after=10 ** 7,
                                            first=None)
    s.close()
    psolve = (f"one multigrid V-cycle for p' (mg_over {c.mg_over:g}"
              + (", chosen by ifl_autotune(IFL_TUNE_PSOLVE_TARGET))" if chosen else ")")
              if c.p_solver == _abi.PSOLVE_MG else "20 p' sweeps")
    return {"workload": f"3D lid-driven cavity {n}^3, Re=400, FP32 RB-GS 10/10/10 momentum sweeps, {psolve}, "
                        "alpha 0.7/0.3",
            "command": f"python3 bench.py --c3-converge --c3-converge-budget-s {args.c3_converge_budget_s:g} "
                       + (f"--c3-mg-over {args.c3_mg_over:g} " if args.c3_mg_over != 0 else "")
                       + "--no-converge --no-mg --no-cpu --no-batch --no-adjoint --no-energy --no-channel --steps 5 --warmup 2",
            "targets": {f"{t:g}": ({"iterations": hit[t][0], "device_s": hit[t][1]} if hit[t] else None)
                        for t in targets},
            "min_scaled": best, "iterations_run": done, "device_s": dev_s}


def adjoint_leg(args, incompflow, _abi):
    """The 1024x256 channel (C5): forward solve from rest, the discrete adjoint of the pressure drop
    (stopped when dJ/dmu and the profile gradient change by less than 1e-6 over 500 iterations), and
    central differences of FP32 device solves warm-started from the converged state file, along mu
    and along the profile itself."""
    import tempfile

    import numpy as np
    c = _abi.adjoint_channel_config(1024, 256)
    c.stop_rule = _abi.STOP_SCALED
    c.tol_residual = args.adjoint_forward_target
    c.num_outer_iter = 200000
    c.sweep_block = args.sweep_block_batch
    prof = np.array(_abi.adjoint_inlet_profile(c))
    launch = None
    if not args.no_tune:
        # the forward's launch settings by the autotuner on a scratch handle of the same case (they
        # change the time per iteration and where a met smoother stop test takes effect, not the
        # converged state; the solve below starts from rest)
        t = incompflow.Solver(c)
        t.set_face_profile(_abi.WEST, prof)
        launch = t.autotune(_abi.TUNE_LAUNCH)
        t.close()
        c.tile, c.sweep_block, c.sweep_block_mom = launch["tile"], launch["sweep_block"], launch["sweep_block_mom"]
        log(f"[bench] adjoint channel launch settings by the autotuner: tile {c.tile}, sweep_block "
            f"{c.sweep_block}, sweep_block_mom {c.sweep_block_mom}")
    log(f"[bench] adjoint channel 1024x256: forward to {c.tol_residual:g}")
    s = incompflow.Solver(c)
    s.set_face_profile(_abi.WEST, prof)
    device_sync()
    t0 = time.perf_counter()
    rf = s.solve(0, history=False)
    device_sync()
    tf = time.perf_counter() - t0
    log(f"[bench]   forward {rf['iterations']} iterations, {tf:.2f} s, converged {rf['converged']}")
    tmp = tempfile.mkdtemp(prefix="ifl_bench_")
    base_state = os.path.join(tmp, "base.state")
    s.save_state(base_state)
    t0 = time.perf_counter()
    ra = s.adjoint(_abi.OBJ_PRESSURE_DROP, _abi.WEST, tol=1e-7, grad_tol=1e-6, check_every=500, max_iter=200000)
    device_sync()
    ta = time.perf_counter() - t0
    s.close()
    J, gmu, gp = ra["J"], ra["dJ_dmu"], ra["dJ_dprofile"]
    log(f"[bench]   adjoint {ra['iterations']} iterations, {ta:.2f} s, J {J:.6e}, dJ/dmu {gmu:.6e}")
    out = {"workload": "2D channel 1024x256 on [0,0.5]x[0,0.05], rho 1, mu 1e-3, parabolic inlet "
                       "(u_max 0.02) + 1e-3 SplitMix64/Box-Muller noise, J = mean p(i=0) - mean p(i=nx-1)",
           "forward_target": c.tol_residual, "forward_target_requested": 1e-6,
           "forward_target_note": "the workload names 1e-6; the FP32 forward solve stalls above it at "
                                  "this size, so the forward stops at 1e-5 (its error against FP64 at "
                                  "256x64 is in fp64_error)",
           "forward_iterations": rf["iterations"], "forward_converged": rf["converged"],
           "forward_final_scaled": rf["final_scaled"], "forward_s": tf,
           "forward_launch_settings": ({"tile": c.tile, "sweep_block": c.sweep_block,
                                        "sweep_block_mom": c.sweep_block_mom, "from": "IFL_TUNE_LAUNCH"}
                                       if launch else None),
           "adjoint_iterations": ra["iterations"], "adjoint_converged": ra["converged"], "adjoint_s": ta,
           "adjoint_grad_change": ra["grad_change"],
           "adjoint_over_forward": ta / tf, "adjoint_ms_per_iteration": ra["elapsed_ms"] / max(1, ra["iterations"]),
           "J": J, "dJ_dmu": gmu, "dJ_dprofile_norm": float(np.linalg.norm(gp)),
           # creeping-flow identities (J linear in mu and in the inflow at Re ~ 1)
           "mu_dJdmu_over_J": c.viscosity * gmu / J, "theta_dJdtheta_over_J": float(np.dot(prof, gp)) / J,
           "fd": None}
    # the gradient against central differences of converged FP64 oracle solves at this size (step 1e-3
    # relative), measured by tests/c5_adjoint_fd.py and kept under profiles/
    ffile = os.path.join(ROOT, "profiles", "adjoint_fd_fp64_c5_1024x256.json")
    if os.path.exists(ffile):
        with open(ffile) as fh:
            ff = json.load(fh)
        out["fd_fp64_oracle"] = {k: ff[k] for k in ("fd_step_relative", "fd_tol", "dJ_dmu_fd", "dJ_dmu_adjoint",
                                                     "dJ_dmu_rel_err", "dir_profile_fd", "dir_profile_adjoint",
                                                     "dir_profile_rel_err", "J_oracle_fp64", "J_device_fp32")}
        out["fd_fp64_oracle"]["source"] = os.path.relpath(ffile, ROOT)
        out["fd_fp64_oracle"]["this_run"] = {
            "dJ_dmu_rel_err": abs(gmu - ff["dJ_dmu_fd"]) / abs(ff["dJ_dmu_fd"]),
            "dir_profile_rel_err": abs(float(np.dot(prof, gp)) - ff["dir_profile_fd"]) / abs(ff["dir_profile_fd"])}
        log(f"[bench]   against FP64-oracle differences: dJ/dmu rel err "
            f"{out['fd_fp64_oracle']['this_run']['dJ_dmu_rel_err']:.2e}, profile direction rel err "
            f"{out['fd_fp64_oracle']['this_run']['dir_profile_rel_err']:.2e}")
    if args.adjoint_fd32:
        fd_its = []

        def J_of(mu, p):
            c2 = c.copy()
            c2.viscosity = mu
            c2.num_outer_iter = 60000          # a solve that stalls at its FP32 floor stops here
            s2 = incompflow.Solver(c2)
            s2.set_face_profile(_abi.WEST, p)
            s2.load_state(base_state)        # warm start from the converged base state
            r2 = s2.solve(0, history=False)
            fd_its.append(r2["iterations"])
            pp = s2.get(_abi.FIELD_P).astype(np.float64)
            s2.close()
            return float(pp[:, :, 0].mean() - pp[:, :, c.nx - 1].mean())
        h = args.adjoint_fd_step
        t0 = time.perf_counter()
        fd_mu = (J_of(c.viscosity * (1 + h), prof) - J_of(c.viscosity * (1 - h), prof)) / (2 * h * c.viscosity)
        fd_dir = (J_of(c.viscosity, prof * (1 + h)) - J_of(c.viscosity, prof * (1 - h))) / (2 * h)
        tfd = time.perf_counter() - t0
        adj_dir = float(np.dot(prof, gp))
        out["fd"] = {"step": h, "dJ_dmu_fd": fd_mu, "dJ_dmu_rel_err": abs(gmu - fd_mu) / abs(fd_mu),
                     "dir_profile_fd": fd_dir, "dir_profile_adjoint": adj_dir,
                     "dir_profile_rel_err": abs(adj_dir - fd_dir) / abs(fd_dir),
                     "solves": 4, "solve_iterations": fd_its, "wall_s": tfd,
                     "kind": "central differences of FP32 device solves to the forward target, warm-started "
                             "from the converged state file"}
        log(f"[bench]   FD check: dJ/dmu rel err {out['fd']['dJ_dmu_rel_err']:.2e}, "
            f"profile direction rel err {out['fd']['dir_profile_rel_err']:.2e} ({tfd:.1f} s)")
    try:
        os.remove(base_state)
        os.rmdir(tmp)
    except OSError:
        pass
    return out


def energy_leg(args, incompflow, _abi, base_ms):
    """The 2D cavity with heated walls (south 373 K, north 293 K, east and west adiabatic, k / c_p =
    mu / 0.71): ms per outer iteration with the temperature equation on (10 sweeps per iteration),
    against the same case without it, and the temperature sweep launch with its algorithmic GB/s."""
    n = args.n
    c = _abi.cavity_config(n=n, p_refine=args.p_refine)
    c.stop_rule = _abi.STOP_NONE
    c.sweep_block = args.sweep_block
    c.sweep_block_mom = args.sweep_block_mom
    c.tile = args.tile
    c.energy, c.n_iter_t, c.relax_t, c.tol_t = 1, 10, 0.7, 1e-5
    c.conductivity, c.specific_heat = c.viscosity / 0.71, 1.0
    for f in range(6):
        c.temp_bc_type[f], c.temp_bc_val[f] = _abi.TEMP_NEUMANN, 0.0
    c.temp_bc_type[_abi.NORTH], c.temp_bc_val[_abi.NORTH] = _abi.TEMP_DIRICHLET, 293.0
    c.temp_bc_type[_abi.SOUTH], c.temp_bc_val[_abi.SOUTH] = _abi.TEMP_DIRICHLET, 373.0
    c.init[4] = 293.0
    s = incompflow.Solver(c)
    s.solve(args.warmup, history=False)
    device_sync()
    t0 = time.perf_counter()
    r = s.solve(args.steps, history=True)
    device_sync()
    ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    us, _ = s.time_kernel(_abi.KERNEL_TEMP_SWEEP, reps=max(10, args.kernel_reps // 4))
    s.close()
    kb = args.sweep_block_mom if args.sweep_block_mom > 0 else min(2, max(1, args.sweep_block))
    cells = n * n
    log(f"[bench] energy: {ms:.4f} ms/step with temperature ({base_ms:.4f} without), T sweep launch {us:.2f} us")
    return {"workload": f"2D cavity {n}x{n}, Re 100, walls south 373 K / north 293 K, east/west adiabatic, "
                        "k/c_p = mu/0.71, 10 temperature sweeps per outer iteration",
            "ms_per_step": ms, "ms_per_step_without": base_ms, "overhead": ms / base_ms - 1.0,
            "sweeps_t_per_step": sum(x.sweeps_t for x in r["history"]) / args.steps,
            "temp_sweep_us": us, "temp_sweep_model_gbs": kb * 32.0 * cells / (us * 1e3),
            "temp_sweep_own_gbs": 20.0 * cells / (us * 1e3)}


def copy_probe():
    """A plain device copy of known size: what the HBM counters read for it scales their readings
    on the solver's kernels."""
    import torch
    n = 64 * 1024 * 1024          # 64 Mi floats = 256 MiB each way
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"[bench] copied 5 x {n * 4} bytes")


def host_cpu():
    """CPU model, sockets, physical cores and logical CPUs of this host (from /proc/cpuinfo)."""
    model, phys, logical = None, set(), 0
    try:
        cur = {}
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if not line.strip():
                    if "processor" in cur:
                        logical += 1
                        phys.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
                if k.strip() == "model name" and model is None:
                    model = v.strip()
        if "processor" in cur:
            logical += 1
            phys.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
    except OSError:
        pass
    sockets = len({p for p, _ in phys}) or None
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = None
    return {"cpu_model": model, "sockets": sockets, "physical_cores": len(phys) or None, "logical_cpus": logical or None,
            "cpus_allowed": allowed, "omp_places": os.environ.get("OMP_PLACES"),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}


def solve_from_rest(incompflow, cfg, budget_s):
    """Solve cfg from rest with its own stop rule in chunks of 5000 outer iterations until it
    converges, fails or the wall budget ends; return (wall seconds, iterations, converged)."""
    c = incompflow.Solver(cfg)
    device_sync()
    t0 = time.perf_counter()
    total = 0
    converged = False
    while True:
        rr = c.solve(5000, history=False)
        total += rr["iterations"]
        log(f"[bench]   {total} iterations, scaled residual {rr['final_scaled']:.3e}, {time.perf_counter() - t0:.1f} s")
        if rr["converged"] or rr["status"] != 0:
            converged = rr["converged"]
            break
        if time.perf_counter() - t0 > budget_s or total >= cfg.num_outer_iter:
            break
    device_sync()
    wall = time.perf_counter() - t0
    c.close()
    return wall, total, converged


def solve_tracking(solver, targets, chunk, budget_s, extra_after, first=0):
    """Run a solver (whose own stop rule never triggers) in chunks with its history; return the
    iteration and device time at which each target scaled residual was first reached, the minimum
    reached, the iterations run and their device time.  After the first target is reached the solve
    goes on for `extra_after` iterations (to observe the floor), or until the wall budget ends."""
    hit = {t: None for t in targets}
    done, best, ms, t0 = 0, float("inf"), 0.0, time.perf_counter()
    stop_at = None
    while stop_at is None or done < stop_at:
        r = solver.solve(chunk if stop_at is None else min(chunk, stop_at - done), history=True)
        h = r["history"]
        per = r["elapsed_ms"] / max(1, len(h))
        for k, rec in enumerate(h):
            for t in targets:
                if hit[t] is None and rec.scaled < t:
                    hit[t] = (done + k + 1, (ms + per * (k + 1)) / 1000.0)
        best = min([best] + [x.scaled for x in h])
        done += r["iterations"]
        ms += r["elapsed_ms"]
        log(f"[bench]   {done} iterations, scaled residual {h[-1].scaled:.3e} (min {best:.3e}), {ms / 1000:.1f} s")
        if stop_at is None and first is not None and hit[targets[first]] is not None:
            stop_at = hit[targets[0]][0] + extra_after
        if r["status"] != 0 or time.perf_counter() - t0 > budget_s:
            break
    return hit, best, done, ms / 1000.0


def mg_leg(args, incompflow, _abi):
    """The 1024^2 cavity with the p' solver the autotuner picks by time to the target on a 4x coarser
    copy of the case (IFL_TUNE_PSOLVE_TARGET; the rate-based IFL_TUNE_PSOLVE choice, from iteration
    --tune-start, is reported beside it: its short trials rank the factors unreliably on this slowly
    converging case), then a solve from rest with it, past the target to the FP32 floor."""
    c = _abi.cavity_config(n=args.n, p_refine=0, tol_residual=1e-12, max_iter=10 ** 7)
    c.sweep_block, c.sweep_block_mom, c.tile = args.sweep_block, args.sweep_block_mom, args.tile
    t = incompflow.Solver(c)
    t.solve(args.tune_start, history=False)
    choice = t.autotune(_abi.TUNE_PSOLVE, trial_iters=args.tune_iters)
    t.close()
    log(f"[bench] p' solver chosen by the autotuner: {choice['p_solver']} (mg_over {choice['mg_over']}, "
        f"{choice['decades_per_s']:.2f} decades/s)")
    ct = c.copy()
    ct.stop_rule, ct.tol_residual = _abi.STOP_SCALED, args.target
    t = incompflow.Solver(ct)
    t0 = time.perf_counter()
    target = t.autotune(_abi.TUNE_PSOLVE_TARGET)
    tune_s = time.perf_counter() - t0
    t.close()
    log(f"[bench] p' solver by time to {args.target:g} on the coarsened cavity: {target['p_solver']} "
        f"(mg_over {target['mg_over']}, {target['trials']} trials, {tune_s:.1f} s)")
    c.p_solver, c.mg_cycles, c.mg_over = target["p_solver"], max(1, target["mg_cycles"]), target["mg_over"]
    if c.p_solver == _abi.PSOLVE_SWEEPS:
        c.p_refine = 1
    s = incompflow.Solver(c)
    s.solve(args.warmup, history=False)
    r = s.solve(args.steps, history=False)
    ms = r["elapsed_ms"] / args.steps
    us, _ = s.time_kernel(_abi.KERNEL_PSOLVE, reps=20)
    s.close()
    s = incompflow.Solver(c)
    log(f"[bench] multigrid p' solve from rest past {args.target:g} (floor: {args.floor_extra} more iterations)")
    hit, best, done, dev_s = solve_tracking(s, [args.target, 1e-5], 10000, args.converge_budget_s,
                                            extra_after=args.floor_extra)
    s.close()
    h4 = hit[args.target]
    out = {"p_solver": "multigrid" if c.p_solver == _abi.PSOLVE_MG else "sweeps", "mg_cycles": c.mg_cycles,
           "mg_over": c.mg_over, "chosen_by": f"ifl_autotune(IFL_TUNE_PSOLVE_TARGET): time to {args.target:g} on "
                                             f"a 4x coarser copy of the cavity, {target['trials']} trials, {tune_s:.1f} s",
           "autotuner_choice": {"p_solver": "multigrid" if choice["p_solver"] == _abi.PSOLVE_MG else "sweeps",
                                "mg_over": choice["mg_over"], "decades_per_s": choice["decades_per_s"],
                                "trials": choice["trials"], "tune_start": args.tune_start, "trial_iters": args.tune_iters},
           "ms_per_step": ms, "p_solve_us": us,
           "iterations_to_residual": h4[0] if h4 else None, "time_to_residual_s": h4[1] if h4 else None,
           "residual_target": args.target, "reached_1e-5": hit[1e-5] is not None,
           "time_to_1e-5_s": hit[1e-5][1] if hit[1e-5] else None,
           "fp32_floor_observed": best, "iterations_run": done}
    log(f"[bench] multigrid: {out['iterations_to_residual']} iterations, {out['time_to_residual_s']} s to "
        f"{args.target:g}; minimum scaled residual {best:.3e} after {done} iterations")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--no-refine-ab", action="store_true",
                    help="skip the solve from rest without the p' refinement cycle")
    ap.add_argument("--target", type=float, default=1e-4)
    ap.add_argument("--converge-budget-s", type=float, default=120.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=600)
    ap.add_argument("--cpu-iters-1thread", type=int, default=12)
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--sweep-block", type=int, default=5,
                    help="p' sweeps per launch on the 2D cavity")
    ap.add_argument("--sweep-block-mom", type=int, default=5,
                    help="momentum/temperature sweeps per launch on the 2D cavity (batch and adjoint legs keep 2)")
    ap.add_argument("--sweep-block-batch", type=int, default=4,
                    help="p' sweeps per launch of the adjoint leg (small tiles: a smaller halo pays)")
    ap.add_argument("--tile", type=int, default=0,
                    help="2D tile (0: the autotuner picks tile and sweeps per launch unless --no-tune)")
    ap.add_argument("--no-tune", action="store_true", help="keep --tile/--sweep-block as given (tile 0: by grid size)")
    ap.add_argument("--p-refine", type=int, default=1)
    ap.add_argument("--mg-over", type=float, default=0.0,
                    help="> 0: the timed workload solves p' by one multigrid V-cycle with this mg_over")
    ap.add_argument("--workload", choices=["cavity2d", "cavity3d"], default="cavity2d")
    ap.add_argument("--no-mg", action="store_true")
    ap.add_argument("--tune-iters", type=int, default=2000)
    ap.add_argument("--tune-start", type=int, default=2000,
                    help="outer iterations run before the p' solver trials (past the start-up transient)")
    ap.add_argument("--floor-extra", type=int, default=60000,
                    help="multigrid iterations past the target, to observe the FP32 floor")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-target", type=float, default=1e-4)
    ap.add_argument("--batch-max-iter", type=int, default=20000)
    ap.add_argument("--batch-sequential", choices=["estimate", "full", "none"], default="full")
    ap.add_argument("--batch-probe-iters", type=int, default=300)
    ap.add_argument("--batch-trial-iters", type=int, default=150)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device to device 5 times (scale for the FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--no-cavity3d", action="store_true")
    ap.add_argument("--no-channel", action="store_true")
    ap.add_argument("--c3-n", type=int, default=256)
    ap.add_argument("--c3-steps", type=int, default=20)
    ap.add_argument("--c3-cpu-iters", type=int, default=50)
    ap.add_argument("--c3-converge", action="store_true",
                    help="solve the 3D cavity from rest towards 1e-5 within --c3-converge-budget-s")
    ap.add_argument("--c3-converge-budget-s", type=float, default=600.0)
    ap.add_argument("--c3-converge-out", default=None, help="write the --c3-converge result to this JSON file")
    ap.add_argument("--c3-mg-over", type=float, default=0.0,
                    help="> 0: the --c3-converge run solves p' by one multigrid V-cycle with this mg_over; "
                         "< 0: with the p' solver the autotuner picks by time to 1e-4 on a coarser copy")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-energy", action="store_true")
    ap.add_argument("--adjoint-forward-target", type=float, default=1e-5)
    ap.add_argument("--adjoint-fd32", action="store_true",
                    help="also difference FP32 device solves (their residual floor limits them to ~1e-3)")
    ap.add_argument("--adjoint-fd-step", type=float, default=0.1,
                    help="relative central-difference step; J is linear in mu and in the inflow up to "
                         "O(Re) terms, so a large step keeps the FP32 solve noise small against the difference")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU slab decomposition is a design, not an implementation.", file=sys.stderr)
        sys.exit(2)

    import incompflow
    from incompflow import _abi

    device_sync()          # torch's own device state, before the timed work (see the import above)
    n = args.n
    if args.workload == "cavity3d":
        if n == 1024:
            n = 256
        cfg = _abi.cavity3d_config(n=n, tol_residual=args.target)
        cfg.p_refine = args.p_refine
    else:
        cfg = _abi.cavity_config(n=n, p_refine=args.p_refine, tol_residual=args.target)
    cfg.sweep_block = args.sweep_block
    cfg.sweep_block_mom = args.sweep_block_mom
    cfg.tile = args.tile
    if args.mg_over > 0:            # the multigrid p' solve instead of the sweeps (profiling runs)
        cfg.p_solver, cfg.mg_cycles, cfg.mg_over, cfg.p_refine = _abi.PSOLVE_MG, 1, args.mg_over, 0
    cells = cfg.nx * cfg.ny * cfg.nz

    # ---- timed steps ----
    run_cfg = cfg.copy()
    run_cfg.stop_rule = _abi.STOP_NONE
    s = incompflow.Solver(run_cfg)
    launch_tune = None
    if not args.no_tune and not args.tile and cfg.smoother == _abi.SMOOTHER_RBGS:
        # the launch settings (2D: tile and sweeps per launch; 3D: p' sweeps per launch) by the
        # library's autotuner, before the warmup; they change the time per iteration (and where a
        # met smoother stop test takes effect: the end of a block), not the converged state
        launch_tune = s.autotune(_abi.TUNE_LAUNCH)
        cfg.tile, cfg.sweep_block, cfg.sweep_block_mom = s.cfg.tile, s.cfg.sweep_block, s.cfg.sweep_block_mom
        args.tile, args.sweep_block, args.sweep_block_mom = cfg.tile, cfg.sweep_block, cfg.sweep_block_mom
        log(f"[bench] launch settings by the autotuner ({launch_tune['trials']} trials): tile {cfg.tile}, "
            f"sweep_block {cfg.sweep_block}, sweep_block_mom {cfg.sweep_block_mom}, "
            f"{launch_tune['us_per_iter']:.1f} us/iteration")
    log(f"[bench] cavity {n}x{n}: warmup {args.warmup} outer iterations")
    s.solve(args.warmup, history=False)
    device_sync()
    log(f"[bench] timing {args.steps} outer iterations")
    t0 = time.perf_counter()
    r = s.solve(args.steps, history=True)
    device_sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    sweeps = sum(sum(rec.sweeps) for rec in r["history"])
    ms_per_step = 1000.0 * wall / args.steps
    value = cells * sweeps / wall
    log(f"[bench] {ms_per_step:.4f} ms/step (device {r['elapsed_ms'] / args.steps:.4f} ms), "
        f"{sweeps / args.steps:.1f} sweeps/step")

    # the whole outer iteration against the byte model of the reference data flow: 38 (2D) / 51 (3D)
    # words per cell for the non-sweep stages plus 8 / 10 words per cell and executed sweep
    words = (38 + 8 * sweeps / args.steps) if cfg.dim == 2 else (51 + 10 * sweeps / args.steps)
    iter_model_bytes = 4.0 * words * cells

    # ---- roofline of the dominant kernel (one launch of the fused red-black p' sweeps) ----
    us, per = s.time_kernel(_abi.KERNEL_SWEEP, reps=args.kernel_reps)
    # sweeps per launch of the p' kernel (3D: one or two)
    kb = max(1, args.sweep_block) if cfg.dim == 2 else min(max(1, args.sweep_block), 2)
    # compulsory bytes of one launch: p' in/out, b' and the coefficient planes towards +x, +y (+z)
    # (those towards -x, -y, -z are the neighbours'), once for all kb sweeps of the launch
    alg_bytes = (20.0 * cells if cfg.dim == 2 else 24.0 * cells) / per
    achieved = alg_bytes / (us * 1e-6) / 1e9
    # the per-sweep streaming model (a stored 5/7-point sweep reads its matrix every sweep)
    model_bytes = kb * sweep_bytes_per_cell(cfg.dim) * cells / per
    # HBM traffic of the same kernel on the same workload, from the committed PMC summary
    # (FETCH_SIZE and WRITE_SIZE runs, scaled by a known-size copy): null on other workloads
    traffic = None
    traffic_source = None
    iteration_traffic = None
    tile = (args.tile or _abi.default_tile(cfg.nx, cfg.ny)) if cfg.dim == 2 else 0
    tname = {_abi.TILE_128x32: "128x32", _abi.TILE_64x16: "64x16"}.get(tile)
    tfile = os.path.join(ROOT, "profiles", f"k_rb_sweep2d_pc_{tname}_cavity1024_k{kb}_traffic.json")
    if cfg.dim == 3:
        tfile = os.path.join(ROOT, "profiles", "k_rb_zmarch3d_pc_32x8x32_cavity256_traffic.json")
    if ((cfg.dim == 2 and n == 1024 and tname) or (cfg.dim == 3 and n == 256 and args.p_refine == 0 and kb == 1)) \
            and os.path.exists(tfile):
        with open(tfile) as fh:
            tj = json.load(fh)
        traffic = tj["traffic_bytes_per_launch"]
        iteration_traffic = tj.get("iteration_traffic_bytes")
        traffic_source = os.path.relpath(tfile, ROOT)
    other = {}
    for name, kid in (("mom_assemble", _abi.KERNEL_MOM_ASSEMBLE), ("rhie_chow", _abi.KERNEL_RHIE_CHOW),
                      ("pcorr_assemble", _abi.KERNEL_PCORR_ASSEMBLE), ("update", _abi.KERNEL_UPDATE),
                      ("momentum_sweep", _abi.KERNEL_MOM_SWEEP), ("p_solve", _abi.KERNEL_PSOLVE)):
        other[name] = round(s.time_kernel(kid, reps=max(10, args.kernel_reps // 4))[0], 3)
    s.close()
    log(f"[bench] sweep launch {us:.3f} us -> {achieved:.1f} GB/s of its compulsory bytes")

    # ---- time to residual (the precision form: sweeps + one refinement cycle) ----
    ttr = None
    iters_to_res = None
    converged = None
    if not args.no_converge:
        log(f"[bench] solving from rest to scaled residual {cfg.tol_residual:g} (budget {args.converge_budget_s:.0f} s)")
        ttr, iters_to_res, converged = solve_from_rest(incompflow, cfg, args.converge_budget_s)

    # ---- the same solve without the refinement cycle: what the FP64-defect refinement buys ----
    refine_ab = None
    if not args.no_converge and not args.no_refine_ab and cfg.dim == 2 and cfg.p_refine > 0 \
            and cfg.p_solver == _abi.PSOLVE_SWEEPS:
        c0 = cfg.copy()
        c0.p_refine = 0
        q = c0.copy()
        q.stop_rule = _abi.STOP_NONE
        s0 = incompflow.Solver(q)
        s0.solve(args.warmup, history=False)
        r0 = s0.solve(args.steps, history=False)
        s0.close()
        ms0 = r0["elapsed_ms"] / args.steps
        log(f"[bench] without refinement: {ms0:.4f} ms/step (device); solving from rest to {c0.tol_residual:g}")
        t0r, it0, conv0 = solve_from_rest(incompflow, c0, args.converge_budget_s)
        refine_ab = {"p_refine_1": {"iterations": iters_to_res, "time_to_residual_s": ttr, "converged": converged,
                                    "ms_per_step_device": r["elapsed_ms"] / args.steps},
                     "p_refine_0": {"iterations": it0, "time_to_residual_s": t0r, "converged": conv0,
                                    "ms_per_step_device": ms0},
                     "residual_target": cfg.tol_residual,
                     "note": "the same 1024^2 cavity and launch settings from rest, with and without the "
                             "FP64-defect refinement cycle of the p' solve (20 more p' sweeps on the defect)"}
        log(f"[bench] refinement A/B to {cfg.tol_residual:g}: with {iters_to_res} iterations {ttr:.2f} s, "
            f"without {it0} iterations {t0r:.2f} s")

    mg = None
    if not args.no_mg and cfg.dim == 2:
        mg = mg_leg(args, incompflow, _abi)

    # error against the FP64 oracle: measured by tests/test_gpu_fp64_error.py (the oracle is the
    # checker there), summarised under profiles/
    fp64 = None
    ffile = os.path.join(ROOT, "profiles", "fp64_error_cavity256_channel256x64.json")
    if cfg.dim == 2 and os.path.exists(ffile):
        with open(ffile) as fh:
            fp64 = json.load(fh)
        fp64["source"] = os.path.relpath(ffile, ROOT)

    # ---- CPU baseline: FP32 oracle on the host cores, bounded sample ----
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        threads = orc.get_threads()
        o = orc.Oracle(cfg)
        it_cpu = args.cpu_iters
        o.solve(1, history=False)   # first iteration (face initialisation) outside the sample
        t0 = time.perf_counter()
        ro = o.solve(it_cpu, history=True)
        tc = time.perf_counter() - t0
        csweeps = sum(sum(rec.sweeps) for rec in ro["history"])
        cpu = {"value": cells * csweeps / tc, "unit": "cell-updates/s", "cores": threads, "kind": "port",
               "sample": f"outer iterations 2..{it_cpu + 1} of the {n}x{n} cavity (FP32 oracle, OpenMP, "
                         f"same settings), {1000 * tc / it_cpu:.2f} ms/iteration",
               "ms_per_step": 1000 * tc / it_cpu, "threads": threads}
        cpu.update(host_cpu())
        if iters_to_res:
            cpu["time_to_residual_s_extrapolated"] = tc / it_cpu * iters_to_res
        if args.cpu_iters_1thread > 0:
            orc.set_threads(1)
            t0 = time.perf_counter()
            r1 = o.solve(args.cpu_iters_1thread, history=True)
            t1c = time.perf_counter() - t0
            orc.set_threads(threads)
            s1 = sum(sum(rec.sweeps) for rec in r1["history"])
            cpu["one_thread"] = {"ms_per_step": 1000 * t1c / args.cpu_iters_1thread, "value": cells * s1 / t1c,
                                 "sample": f"{args.cpu_iters_1thread} further outer iterations on 1 thread"}
        o.close()
        log(f"[bench] cpu oracle {1000 * tc / it_cpu:.1f} ms/iteration on {threads} threads"
            + (f", {cpu['one_thread']['ms_per_step']:.0f} ms on 1 thread" if "one_thread" in cpu else ""))

    batch = None
    if not args.no_batch and cfg.dim == 2:
        batch = batch_leg(args, incompflow, _abi)
    adjoint = None
    if not args.no_adjoint and cfg.dim == 2:
        adjoint = adjoint_leg(args, incompflow, _abi)
    channel = None
    if not args.no_channel and cfg.dim == 2:
        channel = channel_leg(args, incompflow, _abi)
    cavity3d = None
    if not args.no_cavity3d and cfg.dim == 2:
        cavity3d = cavity3d_leg(args, incompflow, _abi)
    energy = None
    if not args.no_energy and cfg.dim == 2:
        energy = energy_leg(args, incompflow, _abi, ms_per_step)

    traffic_gbs = traffic / (us * 1e-6) / 1e9 if traffic else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (fields start at rest)",
        "config": {"workload": (f"2D lid-driven cavity {n}x{n}, Re=100, FP32 RB-GS 10/10/20 sweeps, "
                                f"alpha 0.7/0.3, FP64 reductions, {args.p_refine} FP64-defect p' refinement")
                   if cfg.dim == 2 else
                   (f"3D lid-driven cavity {n}^3, Re=400, FP32 RB-GS 10/10/10/20 sweeps, alpha 0.7/0.3, "
                    f"FP64 reductions"),
                   "cells": cells, "sweeps_per_step": sweeps / args.steps, "sweep_block": kb, "sweep_block_mom": cfg.sweep_block_mom,
                   "launch_settings_from": ("the autotuner (IFL_TUNE_LAUNCH, %d trials)" % launch_tune["trials"])
                                           if launch_tune else "command line",
                   "tile": {_abi.TILE_128x32: "128x32", _abi.TILE_64x16: "64x16", _abi.TILE_64x8: "64x8",
                            _abi.TILE_32x8: "32x8", _abi.TILE_32x4: "32x4"}.get(tile, "3d")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_source,
                     "traffic_gbs": traffic_gbs,
                     "traffic_frac": traffic_gbs / HBM_PEAK_GBS if traffic_gbs else None,
                     "kernel": ("k_rb_sweep2d<COEF_PCORR> (%d fused red-black sweeps of p' per launch)" % kb
                                if cfg.dim == 2 else ("k_rb_zmarch3d_p2 (two red-black sweeps of p' per launch)" if kb == 2
                                                      else "k_rb_zmarch3d_p (one red-black sweep of p' per launch)")),
                     "avg_launch_us": us, "alg_bytes_per_launch": alg_bytes,
                     "sweep_model_bytes_per_launch": model_bytes,
                     "sweep_model_frac": model_bytes / (us * 1e-6) / (HBM_PEAK_GBS * 1e9),
                     "note": ("achieved/frac: the launch's compulsory bytes (field in and out, b' and the +x/+y "
                              "coefficient planes, once for its %d sweeps) over its time; traffic: the bytes the "
                              "HBM counters measured for it; sweep_model: the per-sweep streaming model (32 B per "
                              "cell and sweep), which a launch that fuses sweeps on one staged copy can exceed" % kb),
                     "iteration_model_frac": iter_model_bytes / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
                     "iteration_traffic_bytes": iteration_traffic,
                     "iteration_traffic_frac": (iteration_traffic / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9)
                                                if iteration_traffic else None)},
        "kernel_us": other,
        "time_to_residual_s": ttr,
        "residual_target": cfg.tol_residual,
        "iterations_to_residual": iters_to_res,
        "converged": converged,
        "refine_ab": refine_ab,
        "time_to_1e-5": {"reached": bool(mg and mg["reached_1e-5"]),
                         "time_s": mg["time_to_1e-5_s"] if mg else None,
                         "fp32_floor_observed": mg["fp32_floor_observed"] if mg else None,
                         "how": "the multigrid solve continued past 1e-4 (its minimum scaled residual); the "
                                "precision form's sweeps reach no lower (their floor is higher)"},
        "p_solver_mg": mg,
        "fp64_error": fp64,
        "gpu_vs_cpu_per_step": None,
        "cpu_baseline": cpu,
        "batch": batch,
        "adjoint": adjoint,
        "channel_3000": channel,
        "cavity3d": cavity3d,
        "energy": energy,
    }
    if cpu:
        out["gpu_vs_cpu_per_step"] = cpu["ms_per_step"] / ms_per_step
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
