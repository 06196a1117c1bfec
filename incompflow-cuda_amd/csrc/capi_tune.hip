// capi_tune.hip - the autotuner (ifl_autotune).
//
// Every trial starts from one snapshot of the handle's arena and is undone afterwards.  The
// handle's configuration is the caller's again on every error return (a scope guard restores
// it); on success the chosen settings are written into it.
#include "capi_host.hpp"

#include <algorithm>
#include <vector>

using namespace ifl_host;

namespace {

struct Settings {
    int tile, K, nIter[4], stopRule, Km;
    int pSolver, mgCycles, pRefine;
    double mgOver;
};

Settings settingsOf(const ifl_solver *h) {
    const ifl_config &c = h->cfg;
    Settings x{h->P.dim == 2 ? h->tile : 0, std::max(1, c.sweep_block), {}, c.stop_rule, c.sweep_block_mom,
               c.p_solver, c.mg_cycles, c.p_refine, c.mg_over};
    for (int q = 0; q < 4; ++q) x.nIter[q] = c.n_iter[q];
    return x;
}

void applySettings(ifl_solver *h, const Settings &x) {
    h->cfg.tile = x.tile;
    if (h->P.dim == 2) h->tile = x.tile;
    h->cfg.sweep_block = x.K;
    h->cfg.sweep_block_mom = x.Km;
    for (int q = 0; q < 4; ++q) h->cfg.n_iter[q] = x.nIter[q];
    h->cfg.stop_rule = x.stopRule;
    h->P.stopRule = x.stopRule;
    h->cfg.p_solver = x.pSolver;
    h->cfg.mg_cycles = x.mgCycles;
    h->cfg.mg_over = x.mgOver;
    h->cfg.p_refine = x.pRefine;
    h->graphDirty = true;
}

// n outer iterations from the current state: device time, scaled residual, divergence.  The
// scaled residual is the smallest over the last quarter of the run (at most the history ring): the
// SIMPLE residual does not fall monotonically, and a single end point misranks settings whose
// residuals oscillate by more than their trends differ over a trial.
int trialRun(ifl_solver *h, int n, double &ms, double &scaled, bool &diverged) {
    int st = buildGraph(h);
    if (st != IFL_OK) return st;
    if ((st = armCase(h, h->stream, 0)) != IFL_OK) return st;
    if (h->iter == 0) ifl::launch_init_faces(planOf(h).L);
    IFL_HIP(hipEventRecord(h->ev0, h->stream));
    for (int q = 0; q < n; ++q) IFL_HIP(hipGraphLaunch(h->graph, h->stream));
    IFL_HIP(hipEventRecord(h->ev1, h->stream));
    IFL_HIP(hipEventSynchronize(h->ev1));
    float f = 0;
    IFL_HIP(hipEventElapsedTime(&f, h->ev0, h->ev1));
    ms = f;
    ifl::CaseState s{};
    if ((st = readState(h, s)) != IFL_OK) return st;
    const long long it0 = h->iter;
    h->iter = s.iter;
    scaled = s.lastScaled;
    diverged = s.done == 2 || !std::isfinite(scaled);
    const long long ran = s.iter - it0, tail = std::min<long long>(ran / 4, kHistCap);
    if (!diverged && tail > 1) {
        std::vector<ifl_iter_record> rec((size_t)tail);
        if ((st = copyHistory(h, h->stream, s.iter - tail, s.iter, rec.data())) != IFL_OK) return st;
        IFL_HIP(hipStreamSynchronize(h->stream));
        for (const ifl_iter_record &r : rec)
            if (std::isfinite(r.scaled) && r.scaled < scaled) scaled = r.scaled;
    }
    return IFL_OK;
}

// Restores the configuration the tuner found (and the tile and stop rule the kernels read) unless
// disarmed; the graph is rebuilt on the next solve.
struct ConfigGuard {
    ifl_solver *h;
    ifl_config c0;
    int tile0;
    bool armed = true;
    ~ConfigGuard() {
        if (!armed) return;
        h->cfg = c0;
        h->tile = tile0;
        h->P.stopRule = c0.stop_rule;
        h->graphDirty = true;
    }
};

}  // namespace

// Device time from rest to `tol` of a copy of the case coarsened by f per side with the p' solver
// of x (ms; < 0 when it did not converge within its cap).
static int proxyTime(ifl_solver *h, const Settings &x, int f, double tol, double &ms) {
    ifl_config c = h->cfg;
    c.nx = std::max(4, h->cfg.nx / f);
    c.ny = std::max(4, h->cfg.ny / f);
    if (c.dim == 3) c.nz = std::max(4, h->cfg.nz / f);
    c.stop_rule = IFL_STOP_SCALED;
    c.tol_residual = tol;
    c.num_outer_iter = 200000;
    c.fail_if_not_converged = 0;
    c.log_every = 0;
    c.tile = 0;
    c.sweep_block = x.K;
    c.sweep_block_mom = x.Km;
    for (int q = 0; q < 4; ++q) c.n_iter[q] = x.nIter[q];
    c.p_solver = x.pSolver;
    c.mg_cycles = x.mgCycles;
    c.mg_over = x.mgOver;
    c.p_refine = x.pRefine;
    ifl_solver *p = nullptr;
    int st = ifl_create(&c, &p);
    if (st != IFL_OK) return st;
    // inlet profiles: each proxy face value is the mean of the f (x f) fine values it covers
    for (int face = 0; face < 2 * c.dim && st == IFL_OK; ++face) {
        if (!h->prof[face]) continue;
        const int a = face < 2 ? h->cfg.ny : h->cfg.nx, b = face < 4 ? h->cfg.nz : h->cfg.ny;
        const int ca = face < 2 ? c.ny : c.nx, cb = face < 4 ? c.nz : c.ny;
        std::vector<float> fine((size_t)a * b);
        if (hipMemcpy(fine.data(), h->prof[face], fine.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            st = fail(IFL_ERR_DEVICE, "reading an inlet profile");
            break;
        }
        std::vector<double> coarse((size_t)ca * cb, 0.0);
        for (int jb = 0; jb < cb; ++jb)
            for (int ja = 0; ja < ca; ++ja) {
                double sum = 0.0;
                int cnt = 0;
                for (int qb = jb * b / cb; qb < (jb + 1) * b / cb; ++qb)
                    for (int qa = ja * a / ca; qa < (ja + 1) * a / ca; ++qa) {
                        sum += fine[(size_t)qa + (size_t)a * qb];
                        ++cnt;
                    }
                coarse[(size_t)ja + (size_t)ca * jb] = cnt ? sum / cnt : 0.0;
            }
        st = ifl_set_face_profile(p, face, coarse.data(), coarse.size());
    }
    ifl_report rep{};
    if (st == IFL_OK) st = ifl_solve(p, 0, &rep, nullptr, 0);
    ifl_destroy(p);
    if (st == IFL_ERR_DIVERGED) st = IFL_OK;
    ms = (st == IFL_OK && rep.converged) ? rep.elapsed_ms : -1.0;
    return st;
}

extern "C" int ifl_autotune(ifl_solver *h, int what, int trial_iters, ifl_tune_result *out) {
    if (!h || !out || what == 0 ||
        (what & ~(IFL_TUNE_LAUNCH | IFL_TUNE_SWEEPS | IFL_TUNE_PSOLVE | IFL_TUNE_PSOLVE_TARGET)))
        return fail(IFL_ERR_ARG, "bad argument");
    std::memset(out, 0, sizeof(*out));
    const int trial = trial_iters > 0 ? trial_iters : 300;
    ConfigGuard guard{h, h->cfg, h->tile};
    const ifl_config c0 = h->cfg;
    Settings best = settingsOf(h);
    best.stopRule = IFL_STOP_NONE;
    int st;
    double ms = 0, sc = 0;
    bool div = false;
    applySettings(h, best);
    // trials on the handle itself (all modes but IFL_TUNE_PSOLVE_TARGET) start from one snapshot
    const bool onHandle = (what & (IFL_TUNE_LAUNCH | IFL_TUNE_SWEEPS | IFL_TUNE_PSOLVE)) != 0;
    if (onHandle && h->iter < 20 && (st = trialRun(h, (int)(20 - h->iter), ms, sc, div)) != IFL_OK) return st;
    Snapshot snap;
    if (onHandle && (st = snapTake(h, snap)) != IFL_OK) return st;
    ifl::CaseState s0{};
    if ((st = readState(h, s0)) != IFL_OK) return st;
    double bestUs = 0.0, bestRate = 0.0;
    auto timeIter = [&](const Settings &x, double &us) -> int {
        applySettings(h, x);
        int r = trialRun(h, 3, ms, sc, div);                      // graph build and warm-up
        if (r == IFL_OK) r = trialRun(h, 30, ms, sc, div);
        if (r == IFL_OK) r = snapRestore(h, snap);
        us = 1000.0 * ms / 30.0;
        out->trials++;
        return r;
    };
    // the drop of the scaled residual per second over `trial` outer iterations (-inf: diverged)
    auto rateOf = [&](const Settings &x, double &rate) -> int {
        applySettings(h, x);
        int r = trialRun(h, trial, ms, sc, div);
        if (r == IFL_OK) r = snapRestore(h, snap);
        out->trials++;
        rate = (div || !(sc > 0.0) || !(s0.lastScaled > 0.0)) ? -1e300
                                                              : std::log10(s0.lastScaled / sc) / (ms / 1000.0);
        return r;
    };
    if ((what & IFL_TUNE_LAUNCH) && h->P.dim == 2 && c0.smoother == IFL_SMOOTHER_RBGS) {
        bestUs = 1e300;
        Settings x = best;
        for (int tile = ifl::TILE_64x16; tile <= ifl::TILE_128x32; ++tile)
            for (int K : {1, 2, 4, 5})
                for (int Km : {0, 4, 5}) {       // momentum: the default rule, or as many as p'
                    if (Km != 0 && Km != K) continue;
                    x.tile = tile;
                    x.K = K;
                    x.Km = Km;
                    double us;
                    if ((st = timeIter(x, us)) != IFL_OK) return st;
                    if (us < bestUs) {
                        bestUs = us;
                        best.tile = tile;
                        best.K = K;
                        best.Km = Km;
                    }
                }
    } else if ((what & IFL_TUNE_LAUNCH) && h->P.dim == 3 && c0.smoother == IFL_SMOOTHER_RBGS) {
        bestUs = 1e300;              // 3D: one or two p' sweeps per launch
        Settings x = best;
        for (int K : {1, 2}) {
            x.K = K;
            double us;
            if ((st = timeIter(x, us)) != IFL_OK) return st;
            if (us < bestUs) {
                bestUs = us;
                best.K = K;
            }
        }
    }
    if (what & IFL_TUNE_SWEEPS) {
        const int uvOpts[3] = {5, 10, 20}, pOpts[3] = {10, 20, 40};
        bestRate = -1e300;
        Settings chosen = best;
        for (int a : uvOpts)
            for (int b : pOpts) {
                Settings x = best;
                x.nIter[0] = x.nIter[1] = x.nIter[2] = a;
                if (h->P.dim == 2) x.nIter[2] = c0.n_iter[2];
                x.nIter[3] = b;
                if (x.pSolver == IFL_PSOLVE_MG && b != pOpts[0]) continue;   // MG ignores the p' sweeps
                double rate;
                if ((st = rateOf(x, rate)) != IFL_OK) return st;
                if (rate > bestRate) {
                    bestRate = rate;
                    chosen = x;
                }
            }
        if (chosen.pSolver == IFL_PSOLVE_MG) chosen.nIter[3] = c0.n_iter[3];
        best = chosen;
        if (bestRate == -1e300) bestRate = 0.0;
    }
    if (what & IFL_TUNE_PSOLVE_TARGET) {
        // the candidates of IFL_TUNE_PSOLVE, by time to the target on the coarsened copy
        const int minSide = std::min(h->cfg.nx, std::min(h->cfg.ny, h->P.dim == 3 ? h->cfg.nz : h->cfg.ny));
        const int f = minSide >= (h->P.dim == 3 ? 96 : 128) ? 4 : minSide >= 64 ? 2 : 1;
        const double tol = c0.stop_rule == IFL_STOP_SCALED && c0.tol_residual > 0 ? c0.tol_residual : 1e-4;
        Settings x = best;
        x.pSolver = IFL_PSOLVE_SWEEPS;
        if (x.pRefine == 0) x.pRefine = c0.p_refine;
        double ms0;
        if ((st = proxyTime(h, x, f, tol, ms0)) != IFL_OK) return st;
        out->trials++;
        Settings chosen = x;
        for (double over : {1.0, 1.4, 1.8}) {
            Settings y = best;
            y.pSolver = IFL_PSOLVE_MG;
            y.pRefine = 0;
            y.mgCycles = 1;
            y.mgOver = over;
            double ms;
            if ((st = proxyTime(h, y, f, tol, ms)) != IFL_OK) return st;
            out->trials++;
            if (ms > 0.0 && (ms0 < 0.0 || ms < ms0)) {
                ms0 = ms;
                chosen = y;
            }
        }
        if (ms0 > 0.0) best = chosen;   // nothing converged: keep the handle's p' solver
        bestRate = 0.0;
    } else if (what & IFL_TUNE_PSOLVE) {
        // the p' solver and its strength: the sweeps as configured (with their refinement cycles),
        // or one multigrid V-cycle with a coarse-correction factor of 1.0, 1.4 or 1.8
        Settings x = best;
        x.pSolver = IFL_PSOLVE_SWEEPS;
        if (x.pRefine == 0) x.pRefine = c0.p_refine;
        double r0;
        if ((st = rateOf(x, r0)) != IFL_OK) return st;
        Settings chosen = x;
        for (double over : {1.0, 1.4, 1.8}) {
            x = best;
            x.pSolver = IFL_PSOLVE_MG;
            x.pRefine = 0;
            x.mgCycles = 1;
            x.mgOver = over;
            double rate;
            if ((st = rateOf(x, rate)) != IFL_OK) return st;
            if (rate > r0) {
                r0 = rate;
                chosen = x;
            }
        }
        best = chosen;
        bestRate = r0 == -1e300 ? 0.0 : r0;
    }
    if (onHandle && (!(what & IFL_TUNE_LAUNCH) || bestUs == 0.0 ||
                     (what & (IFL_TUNE_SWEEPS | IFL_TUNE_PSOLVE | IFL_TUNE_PSOLVE_TARGET))))
        if ((st = timeIter(best, bestUs)) != IFL_OK) return st;
    best.stopRule = c0.stop_rule;
    applySettings(h, best);
    guard.armed = false;
    out->tile = best.tile;
    out->sweep_block = best.K;
    out->sweep_block_mom = best.Km;
    for (int q = 0; q < 4; ++q) out->n_iter[q] = best.nIter[q];
    out->us_per_iter = bestUs;
    out->decades_per_s = bestRate;
    out->p_solver = best.pSolver;
    out->mg_cycles = best.mgCycles;
    out->mg_over = best.mgOver;
    return IFL_OK;
}
