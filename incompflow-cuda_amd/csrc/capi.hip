// capi.hip - host side of the C-ABI in include/incompflow.h (C++20 + HIP runtime).
//
// A handle owns one device arena (all fields, coefficient planes, reduction partials, the
// CaseState monitor and a ring of per-iteration records) and one HIP stream.  One SIMPLE outer
// iteration (Solver.cu:160-290) is captured once into a hipGraph and replayed; the device-side
// stop flag makes every kernel of an already finished case return at once, so the host reads the
// monitor only every `check_every` iterations instead of after every sweep (Solver.cu:951-953)
// and every iteration (Solver.cu:207-224).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "../../include/incompflow.h"
#include "ifl_adjoint.hpp"
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define IFL_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(IFL_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

constexpr int kHistCap = 4096;
constexpr int kMaxSweeps = 100000;

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

struct ifl_solver {
    ifl_config cfg{};
    ifl::CaseParams P{};
    hipStream_t stream = nullptr;
    char *arena = nullptr;
    size_t arenaBytes = 0;
    float *prof[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t graph = nullptr;
    bool graphDirty = true;
    long long iter = 0;   // host mirror of CaseState::iter
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float tolf[5];
    int tile = 0;     // 2D sweep tile variant
    struct AdjHost *adj = nullptr;   // discrete adjoint buffers, made by the first ifl_adjoint call
};

// Device buffers of the adjoint of one handle and the graph of one adjoint iteration.
struct AdjHost {
    char *arena = nullptr;
    ifl::AdjBufs B{};
    float *gw[4] = {nullptr, nullptr, nullptr, nullptr};   // objective weights u, v, w, p
    hipGraphExec_t graph = nullptr;
};

namespace {

int validate(const ifl_config &c) {
    if (c.dim != 2 && c.dim != 3) return fail(IFL_ERR_ARG, "dim must be 2 or 3");
    if (c.nx < 4 || c.ny < 4) return fail(IFL_ERR_ARG, "nx and ny must be >= 4");
    if (c.dim == 2 && c.nz != 1) return fail(IFL_ERR_ARG, "nz must be 1 in 2D");
    if (c.dim == 3 && c.nz < 4) return fail(IFL_ERR_ARG, "nz must be >= 4 in 3D");
    if ((long long)c.nx * c.ny * c.nz > (1LL << 31) / 8) return fail(IFL_ERR_ARG, "grid too large");
    if (!(c.xmax > c.xmin) || !(c.ymax > c.ymin) || !(c.zmax > c.zmin)) return fail(IFL_ERR_ARG, "empty domain");
    for (int f = 0; f < 6; ++f)
        if (c.bc_type[f] < 0 || c.bc_type[f] > 2) return fail(IFL_ERR_ARG, "bc_type must be 0 (wall), 1 (inlet) or 2 (outlet)");
    for (int q = 0; q < 4; ++q)
        if (c.n_iter[q] < 0 || c.n_iter[q] > kMaxSweeps) return fail(IFL_ERR_ARG, "n_iter out of range");
    if (c.n_iter[3] < 1) return fail(IFL_ERR_ARG, "n_iter[p] must be >= 1");
    if (c.smoother != IFL_SMOOTHER_RBGS && c.smoother != IFL_SMOOTHER_JACOBI)
        return fail(IFL_ERR_ARG, "smoother must be IFL_SMOOTHER_RBGS or IFL_SMOOTHER_JACOBI on the device");
    if (c.stop_rule < 0 || c.stop_rule > 3) return fail(IFL_ERR_ARG, "bad stop_rule");
    if (c.p_refine < 0 || c.p_refine > 16) return fail(IFL_ERR_ARG, "p_refine out of range");
    if (c.tile < 0 || c.tile > ifl::TILE_128x32) return fail(IFL_ERR_ARG, "tile must be 0..5");
    if (c.sweep_block < 0 || c.sweep_block > 5 || c.sweep_block == 3)
        return fail(IFL_ERR_ARG, "sweep_block must be 0, 1, 2, 4 or 5");
    if (c.sweep_block_mom < 0 || c.sweep_block_mom > 5 || c.sweep_block_mom == 3)
        return fail(IFL_ERR_ARG, "sweep_block_mom must be 0, 1, 2, 4 or 5");
    if (c.energy != 0 && c.energy != 1) return fail(IFL_ERR_ARG, "energy must be 0 or 1");
    if (c.energy) {
        if (c.n_iter_t < 0 || c.n_iter_t > kMaxSweeps) return fail(IFL_ERR_ARG, "n_iter_t out of range");
        if (!(c.conductivity >= 0) || !(c.specific_heat > 0))
            return fail(IFL_ERR_ARG, "energy needs conductivity >= 0 and specific_heat > 0");
        for (int f = 0; f < 6; ++f)
            if (c.temp_bc_type[f] != IFL_TEMP_DIRICHLET && c.temp_bc_type[f] != IFL_TEMP_NEUMANN)
                return fail(IFL_ERR_ARG, "temp_bc_type must be 0 (Dirichlet) or 1 (Neumann)");
    }
    return IFL_OK;
}

void fillParams(ifl_solver *h) {
    const ifl_config &c = h->cfg;
    ifl::CaseParams &P = h->P;
    P.dim = c.dim; P.nx = c.nx; P.ny = c.ny; P.nz = c.nz;
    P.N = (long long)c.nx * c.ny * c.nz;
    // constants.h:22-28 in float, exactly as the reference's constexprs
    const float xmin = (float)c.xmin, xmax = (float)c.xmax, ymin = (float)c.ymin, ymax = (float)c.ymax;
    const float zmin = (float)c.zmin, zmax = (float)c.zmax;
    P.dx = (xmax - xmin) / (float)c.nx;
    P.dy = (ymax - ymin) / (float)c.ny;
    P.dz = (zmax - zmin) / (float)c.nz;
    P.Ax = P.dy * P.dz;
    P.Ay = P.dz * P.dx;
    P.Az = P.dx * P.dy;
    P.rho = (float)c.density;
    P.mu = (float)c.viscosity;
    for (int f = 0; f < 6; ++f) {
        P.bct[f] = c.bc_type[f];
        for (int q = 0; q < 3; ++q) P.bcv[f][q] = (float)c.bc_val[f][q];
        P.prof[f] = h->prof[f];
    }
    for (int q = 0; q < 4; ++q) {
        P.relax[q] = (float)c.relax[q];
        h->tolf[q] = (float)c.tol[q];
        P.tol[q] = (double)h->tolf[q];
    }
    P.relax[4] = (float)c.relax_t;
    h->tolf[4] = (float)c.tol_t;
    P.tol[4] = (double)h->tolf[4];
    P.tolMass = (double)(float)c.tol_mass;
    P.tolRes = c.tol_residual;
    P.stopRule = c.stop_rule;
    P.histCap = kHistCap;
    // temperature (config.h:19-23, 45-63): diffusion coefficient k / c_p in float arithmetic
    P.energy = c.energy;
    P.gam = (float)c.conductivity / (float)c.specific_heat;
    for (int f = 0; f < 6; ++f) {
        P.tbct[f] = c.temp_bc_type[f];
        P.tbcv[f] = (float)c.temp_bc_val[f];
    }
}

// Records per (slot, launch parity, sweep) partial array: one per workgroup of the largest sweep
// launch (fused 2D tiles: the smallest tile gives the most; 3D z-march; Jacobi cell grid).
int partStride(const ifl::CaseParams &P) {
    int n = std::max({ifl::cell_blocks(P), ifl::zmarch_blocks(P, false), ifl::zmarch_blocks(P, true)});
    if (P.dim == 2)
        for (int t = ifl::TILE_64x16; t <= ifl::TILE_128x32; ++t) n = std::max(n, ifl::tile_count(P.nx, P.ny, t));
    return n;
}

// What one outer iteration launches, for one case or a batch of cases with the same settings.
struct Plan {
    ifl::Lc L;
    const ifl_config *cfg;   // sweep counts, smoother, refinement, sweep block (shared by a batch)
    int dim;
};

// Sweeps per launch of a smoother call in 2D red-black (1 otherwise): sweep_block for p' and its
// refinement, sweep_block_mom for the momentum and temperature equations (0: the smaller of
// sweep_block and 2).
int sweepBlockOf(const ifl_config &c, int dim, int slot) {
    if (dim != 2 || c.smoother != IFL_SMOOTHER_RBGS) return 1;
    if (slot < 3 || slot == ifl::kSlotT) {
        const int k = c.sweep_block_mom > 0 ? c.sweep_block_mom : std::min(std::max(c.sweep_block, 1), 2);
        return std::max(k, 1);
    }
    return std::max(c.sweep_block, 1);
}

// Equation index of a smoother slot (relax / tol / sweep counts): u, v, w, p' (p' and its
// refinement correction), T.
int slotEq(int slot) { return slot < 3 ? slot : slot == ifl::kSlotT ? 4 : 3; }
int slotSweeps(const ifl_config &c, int slot) { return slot == ifl::kSlotT ? c.n_iter_t : c.n_iter[slotEq(slot)]; }

// Workgroup kind of a smoother launch and, for one case, its workgroup count.
int sweepKind(const Plan &pl, int slot) {
    if (pl.cfg->smoother == IFL_SMOOTHER_JACOBI) return ifl::MAP_CELL;
    if (pl.dim == 2) return ifl::MAP_TILE;
    return (slot == 3 || slot == 4) ? ifl::MAP_ZMARCH_P : ifl::MAP_ZMARCH;
}
int sweepRecordsSingle(const Plan &pl, int slot) {
    const ifl::CaseParams &P = *pl.L.P;
    switch (sweepKind(pl, slot)) {
        case ifl::MAP_CELL: return ifl::cell_blocks(P);
        case ifl::MAP_TILE: return ifl::tile_count(P.nx, P.ny, pl.L.tile);
        case ifl::MAP_ZMARCH_P: return ifl::zmarch_blocks(P, true);
        default: return ifl::zmarch_blocks(P, false);
    }
}

// How a smoother call is closed.  The outer-iteration path (fast) lets a call with an odd launch
// count start in the partner buffer (which the kernels keep equal to the field: the update writes
// the velocities to both, p' and its correction start from zero in both), so the result lands in
// the caller's field without a copy, and hands the call's closing step to workgroup 0 of the next
// kernel (pending).  The stage-by-stage path copies back and launches k_smooth_finish, and leaves
// the momentum fields' partners equal to them.
struct SmoothMode {
    bool fast = false;
    ifl::Finish *pending = nullptr;
    bool accumLast = false;      // the last launch adds the correction into p' (2D red-black)
};

// Smoother calls of Solver.cu:178-196.  Every launch reads one buffer of a ping-pong pair and
// writes the other.  Equations listed together share the matrix and are swept in the same
// launches (u and v in 2D); each keeps its own norms and stop state.
void enqueueSmoothers(const Plan &pl, const int *slots, int neq, const SmoothMode &md = SmoothMode{}) {
    const ifl_config &c = *pl.cfg;
    const int K = sweepBlockOf(c, pl.dim, slots[0]);
    const bool fused = c.smoother == IFL_SMOOTHER_RBGS;      // the equations of slots[] in one launch
    if (!fused && neq > 1) {
        for (int k = 0; k < neq; ++k) enqueueSmoothers(pl, &slots[k], 1, md);
        return;
    }
    int n[3], nL[3], flip[3], nLmax = 0;
    for (int k = 0; k < neq; ++k) {
        n[k] = slotSweeps(c, slots[k]);
        nL[k] = (n[k] + K - 1) / K;
        // the temperature has no stage that keeps its partner buffer equal to it: it never starts
        // there, and an odd launch count is copied back
        flip[k] = (md.fast && nL[k] % 2 == 1 && slots[k] != ifl::kSlotT) ? 1 : 0;
        nLmax = std::max(nLmax, nL[k]);
    }
    for (int L = 0; L < nLmax; ++L) {
        ifl::SweepSel se[3];
        int m = 0;
        for (int k = 0; k < neq; ++k) {
            const int kk = std::min(K, n[k] - L * K);
            if (kk <= 0) continue;
            const int kp = L > 0 ? std::min(K, n[k] - (L - 1) * K) : 0;
            se[m] = ifl::SweepSel{slots[k], slotEq(slots[k]), kk, kp};
            se[m].flip = flip[k];
            se[m].accum = (md.accumLast && fused && L == nL[k] - 1) ? 1 : 0;
            ++m;
        }
        if (m == 1) se[1] = se[0];
        if (fused) {
            const int kind = slots[0] < 3 ? ifl::COEF_MOMENTUM
                             : slots[0] == ifl::kSlotT ? ifl::COEF_TEMP : ifl::COEF_PCORR;
            if (pl.dim == 2) ifl::launch_rb_sweep2d(pl.L, kind, se[0], se[1], m, L, K);
            else ifl::launch_rb_sweep3d(pl.L, kind, se, m, L);
        } else {
            for (int k = 0; k < m; ++k) ifl::launch_sweep_jacobi(pl.L, se[k], L);
        }
    }
    for (int k = 0; k < neq; ++k) {
        if (!md.fast || slots[k] == ifl::kSlotT) {
            if (nL[k] % 2 == 1) ifl::launch_copy_back(pl.L, slots[k], 0);
            else if (slots[k] < 3 && nL[k] > 0) ifl::launch_copy_back(pl.L, slots[k], 1);
        }
        if (n[k] > 0) {
            const int q = slotEq(slots[k]);
            const int kLast = n[k] - (nL[k] - 1) * K;
            if (md.fast && md.pending) {
                ifl::Finish &f = *md.pending;
                f.kind = sweepKind(pl, slots[k]);
                f.nprev = sweepRecordsSingle(pl, slots[k]);
                f.slot[f.n] = slots[k];
                f.q[f.n] = q;
                f.nL[f.n] = nL[k];
                f.kLast[f.n] = kLast;
                ++f.n;
            } else {
                ifl::launch_smooth_finish(pl.L, slots[k], q, nL[k], kLast, sweepKind(pl, slots[k]),
                                          sweepRecordsSingle(pl, slots[k]));
            }
        }
    }
}

void enqueueMomentumSmoothers(const Plan &pl, const SmoothMode &md = SmoothMode{}) {
    if (pl.cfg->smoother == IFL_SMOOTHER_RBGS) {      // u, v (and w) share the matrix: one sweep kernel
        const int slots[3] = {0, 1, 2};
        enqueueSmoothers(pl, slots, pl.dim, md);
        return;
    }
    for (int q = 0; q < pl.dim; ++q) enqueueSmoothers(pl, &q, 1, md);   // :178, :180, :183
}

// The temperature sweeps (energy on), after the momentum sweeps on the same face velocities.
void enqueueTemperatureSmoother(const Plan &pl, const SmoothMode &md = SmoothMode{}) {
    if (!pl.cfg->energy) return;
    const int slot = ifl::kSlotT;
    enqueueSmoothers(pl, &slot, 1, md);
}

// Solver.cu:196 plus the FP64-defect refinement cycles.  Returns the calls still to be closed
// (fast path) by the next kernel.
ifl::Finish enqueuePressureCorrection(const Plan &pl, bool fast = false) {
    const int sp = 3, sd = 4;
    const bool fused = pl.dim == 2 && pl.cfg->smoother == IFL_SMOOTHER_RBGS;
    ifl::Finish pend;
    SmoothMode md;
    md.fast = fast;
    md.pending = &pend;
    enqueueSmoothers(pl, &sp, 1, md);
    for (int r = 0; r < pl.cfg->p_refine; ++r) {
        ifl::launch_pcorr_defect(pl.L, pend);
        pend = ifl::Finish{};
        SmoothMode mr = md;
        mr.accumLast = fast && fused && r == pl.cfg->p_refine - 1;
        enqueueSmoothers(pl, &sd, 1, mr);
        if (!mr.accumLast) ifl::launch_pcorr_accumulate(pl.L);
    }
    return pend;
}

void enqueueOuter(const Plan &pl) {
    ifl::launch_mom_assemble(pl.L);                                  // :165-175
    ifl::Finish fm;
    SmoothMode md;
    md.fast = true;
    md.pending = &fm;
    enqueueMomentumSmoothers(pl, md);                                // :178-184
    enqueueTemperatureSmoother(pl, md);
    ifl::launch_rhie_chow(pl.L, fm);                                 // :186-187
    ifl::launch_pcorr_assemble(pl.L);                                // :189-193
    const ifl::Finish fp = enqueuePressureCorrection(pl, true);      // :196
    ifl::launch_update(pl.L, fp);                                    // :198
    ifl::launch_reduce_cells(pl.L, ifl::RED_MOM | ifl::RED_MASS | ifl::RED_UPD | ifl::RED_FINISH);  // :200-268
}

Plan planOf(ifl_solver *h) {
    Plan pl;
    pl.L.P = &h->P;
    pl.L.batch = false;
    pl.L.tile = h->tile ? h->tile : ifl::TILE_64x16;
    pl.L.s = h->stream;
    pl.cfg = &h->cfg;
    pl.dim = h->P.dim;
    return pl;
}

int captureGraph(hipStream_t stream, const Plan &pl, hipGraphExec_t *out) {
    hipGraph_t g = nullptr;
    IFL_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    enqueueOuter(pl);
    IFL_HIP(hipStreamEndCapture(stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(IFL_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return IFL_OK;
}

int buildGraph(ifl_solver *h) {
    if (!h->graphDirty && h->graph) return IFL_OK;
    if (h->graph) {
        IFL_HIP(hipGraphExecDestroy(h->graph));
        h->graph = nullptr;
    }
    const int st = captureGraph(h->stream, planOf(h), &h->graph);
    if (st != IFL_OK) return st;
    h->graphDirty = false;
    return IFL_OK;
}

int readState(ifl_solver *h, ifl::CaseState &s) {
    IFL_HIP(hipMemcpyAsync(&s, h->P.st, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    return IFL_OK;
}

void printLog(const ifl_iter_record &r, long long it, int dim) {
    // Solver.cu:270-285
    std::printf("iter: %lld\n", it + 1);
    std::printf("pRelNorm : %.3E\n", r.rel_p);
    std::printf("uRelNorm : %.3E\n", r.rel_u);
    std::printf("vRelNorm : %.3E\n", r.rel_v);
    if (dim == 3) std::printf("wRelNorm : %.3E\n", r.rel_w);
    std::printf("ufRelNorm: %.3E\n", r.rel_uf);
    std::printf("vfRelNorm: %.3E\n", r.rel_vf);
    if (dim == 3) std::printf("wfRelNorm: %.3E\n", r.rel_wf);
    std::printf("massNorm : %.3E\n", r.mass_norm);
    std::printf("------------------------------------------------------------\n");
}

int setDone(ifl_solver *h, int v) {
    IFL_HIP(hipMemcpyAsync(&h->P.st->done, &v, sizeof(int), hipMemcpyHostToDevice, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    return IFL_OK;
}

// Clear the stop flag and set the iteration cap (0: none) of a case, on stream s.
int armCase(ifl_solver *h, hipStream_t s, long long cap) {
    static const int zero = 0;
    IFL_HIP(hipMemcpyAsync(&h->P.st->done, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    IFL_HIP(hipMemcpyAsync(&h->P.st->iterCap, &cap, sizeof(long long), hipMemcpyHostToDevice, s));
    IFL_HIP(hipStreamSynchronize(s));
    return IFL_OK;
}

// Records of iterations [a, b) from the device ring into dst (b - a <= ring size): two copies at most.
int copyHistory(ifl_solver *h, hipStream_t s, long long a, long long b, ifl_iter_record *dst) {
    const long long n = b - a;
    if (n <= 0) return IFL_OK;
    const long long i0 = a % kHistCap;
    const long long first = std::min<long long>(n, kHistCap - i0);
    IFL_HIP(hipMemcpyAsync(dst, h->P.hist + i0, (size_t)first * sizeof(ifl_iter_record), hipMemcpyDeviceToHost, s));
    if (n > first)
        IFL_HIP(hipMemcpyAsync(dst + first, h->P.hist, (size_t)(n - first) * sizeof(ifl_iter_record),
                               hipMemcpyDeviceToHost, s));
    return IFL_OK;
}

// Host bookkeeping of a solve over the records of iterations [a, b) (it0: first iteration of
// the solve): history output, sweep count, log lines (Solver.cu:270-285).
struct SolveTally {
    long long it0 = 0;
    int maxIter = 0;
    double sweeps = 0.0;
    ifl_iter_record *history = nullptr;
    size_t histCap = 0;
};
void tally(ifl_solver *h, SolveTally &t, const ifl_iter_record *recs, long long a, long long b, bool doneAtEnd,
           int logEvery) {
    for (long long it = a; it < b; ++it) {
        const ifl_iter_record &r = recs[it - a];
        const long long local = it - t.it0;
        if (t.history && (size_t)local < t.histCap) t.history[local] = r;
        for (int q = 0; q < 4; ++q) t.sweeps += (double)r.sweeps[q] * (double)h->P.N;
        t.sweeps += (double)r.sweeps_t * (double)h->P.N;
        const bool last = (local == t.maxIter - 1) || r.converged || (it == b - 1 && doneAtEnd);
        if (logEvery > 0 && (local == 0 || (it + 1) % logEvery == 0 || last)) printLog(r, it, h->P.dim);
    }
}

float *fieldPtr(ifl_solver *h, int id) {
    ifl::CaseParams &P = h->P;
    switch (id) {
        case IFL_FIELD_U: return P.u;
        case IFL_FIELD_V: return P.v;
        case IFL_FIELD_W: return P.w;
        case IFL_FIELD_P: return P.p;
        case IFL_FIELD_TEMP: return P.temp;
        case IFL_FIELD_PCORR: return P.pc;
        case IFL_FIELD_UF: return P.uf;
        case IFL_FIELD_VF: return P.vf;
        case IFL_FIELD_WF: return P.wf;
        case IFL_FIELD_MCOEF: return P.mcoef;
        case IFL_FIELD_PCOEF: return P.pcoef;
        case IFL_FIELD_BU: return P.bu;
        case IFL_FIELD_BV: return P.bv;
        case IFL_FIELD_BW: return P.bw;
        case IFL_FIELD_BP: return P.bp;
        case IFL_FIELD_TCOEF: return P.tcoef;      // null unless energy is on
        case IFL_FIELD_BT: return P.bt;
        default: return nullptr;
    }
}

size_t fieldSize(const ifl_config &c, int id) {
    size_t nx = c.nx, ny = c.ny, nz = c.nz, n = nx * ny * nz;
    switch (id) {
        case IFL_FIELD_UF: return (nx + 1) * ny * nz;
        case IFL_FIELD_VF: return nx * (ny + 1) * nz;
        case IFL_FIELD_WF: return nx * ny * (nz + 1);
        case IFL_FIELD_MCOEF:
        case IFL_FIELD_PCOEF:
        case IFL_FIELD_TCOEF: return n * (1 + 2 * c.dim);
        default: return (id >= 0 && id < IFL_FIELD_COUNT) ? n : 0;
    }
}


}  // namespace

extern "C" {

const char *ifl_last_error(void) { return g_err.c_str(); }
int ifl_abi_version(void) { return IFL_ABI_VERSION; }

void ifl_config_default(ifl_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    // include/config.h:6-81 as shipped
    c->dim = 2;
    c->nx = 1024; c->ny = 256; c->nz = 1;
    c->xmin = -2; c->xmax = 2; c->ymin = -0.5; c->ymax = 0.5; c->zmin = 0; c->zmax = 1;
    c->density = 50; c->viscosity = 1;
    const int types[6] = {IFL_OUTLET, IFL_INLET, IFL_WALL, IFL_WALL, IFL_WALL, IFL_WALL};
    for (int f = 0; f < 6; ++f) c->bc_type[f] = types[f];
    c->bc_val[IFL_WEST][0] = 1;
    c->num_outer_iter = 20000;
    c->n_iter[0] = c->n_iter[1] = c->n_iter[2] = 10;
    c->n_iter[3] = 20;
    c->relax[0] = c->relax[1] = c->relax[2] = 0.75;
    c->relax[3] = 0.25;
    c->tol[0] = c->tol[1] = c->tol[2] = 1e-5;
    c->tol[3] = 1e-3;
    c->tol_mass = 1e-9;
    c->init[4] = 273;   // Solver.h:46 default temp
    // config.h:19-20, 45-63 (declared, not solved by the reference: energy stays 0)
    c->conductivity = 1;
    c->specific_heat = 1;
    const int ttypes[6] = {IFL_TEMP_NEUMANN, IFL_TEMP_NEUMANN, IFL_TEMP_DIRICHLET, IFL_TEMP_DIRICHLET,
                           IFL_TEMP_NEUMANN, IFL_TEMP_NEUMANN};
    const double tvals[6] = {0.0, 0.0, 293.0, 373.0, 0.0, 0.0};
    for (int f = 0; f < 6; ++f) {
        c->temp_bc_type[f] = ttypes[f];
        c->temp_bc_val[f] = tvals[f];
    }
    c->n_iter_t = 10;
    c->relax_t = 0.75;
    c->tol_t = 1e-5;
    c->smoother = IFL_SMOOTHER_RBGS;
    c->stop_rule = IFL_STOP_FAITHFUL;
    c->tol_residual = 1e-5;
    c->log_every = 1000;
}

int ifl_create(const ifl_config *cfg, ifl_solver **out) {
    if (!cfg || !out) return fail(IFL_ERR_ARG, "null argument");
    *out = nullptr;
    int st = validate(*cfg);
    if (st != IFL_OK) return st;
    ifl_solver *h = new (std::nothrow) ifl_solver();
    if (!h) return fail(IFL_ERR_NOMEM, "out of host memory");
    h->cfg = *cfg;
    fillParams(h);
    h->tile = (cfg->dim == 2) ? (cfg->tile ? cfg->tile : ifl::default_tile(cfg->nx, cfg->ny)) : 0;
    ifl::CaseParams &P = h->P;
    const size_t N = (size_t)P.N, nx = P.nx, ny = P.ny, nz = P.nz;
    const size_t ncoef = N * (1 + 2 * P.dim);
    const size_t pstride = (size_t)partStride(P), ncb = (size_t)ifl::cell_blocks(P);
    P.partStride = (int)pstride;
    // arena layout
    struct Part { void **dst; size_t bytes; };
    const size_t nT = cfg->energy ? N : 0;     // temperature matrix, source and partner: energy only
    std::vector<Part> parts = {
        {(void **)&P.u, N * 4}, {(void **)&P.v, N * 4}, {(void **)&P.w, N * 4}, {(void **)&P.p, N * 4},
        {(void **)&P.temp, N * 4}, {(void **)&P.pc, N * 4},
        {(void **)&P.uf, (nx + 1) * ny * nz * 4}, {(void **)&P.vf, nx * (ny + 1) * nz * 4},
        {(void **)&P.wf, nx * ny * (nz + 1) * 4},
        {(void **)&P.mcoef, ncoef * 4}, {(void **)&P.pcoef, ncoef * 4},
        {(void **)&P.bu, N * 4}, {(void **)&P.bv, N * 4}, {(void **)&P.bw, N * 4}, {(void **)&P.bp, N * 4},
        {(void **)&P.tmp, N * 4}, {(void **)&P.tmp2, N * 4},
        {(void **)&P.alt[0], N * 4}, {(void **)&P.alt[1], N * 4}, {(void **)&P.alt[2], N * 4},
        {(void **)&P.alt[3], N * 4}, {(void **)&P.alt[4], N * 4}, {(void **)&P.rinv, N * 4},
        {(void **)&P.tcoef, nT * (1 + 2 * P.dim) * 4}, {(void **)&P.bt, nT * 4}, {(void **)&P.alt[5], nT * 4},
        {(void **)&P.spart, (size_t)ifl::kSlots * 2 * ifl::kMaxBlock * pstride * 8},
        {(void **)&P.partMom, 4 * ncb * 8}, {(void **)&P.partMass, ncb * 8}, {(void **)&P.partUpd, 7 * ncb * 8},
        {(void **)&P.st, sizeof(ifl::CaseState)},
        {(void **)&P.hist, sizeof(ifl_iter_record) * kHistCap},
    };
    size_t total = 0;
    for (auto &p : parts) total += align256(p.bytes);
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&h->arena, total);
    h->arenaBytes = total;
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e != hipSuccess) {
        std::string m = std::string("device allocation: ") + hipGetErrorString(e);
        ifl_destroy(h);
        return fail(IFL_ERR_DEVICE, m);
    }
    size_t off = 0;
    for (auto &p : parts) {
        *p.dst = p.bytes ? h->arena + off : nullptr;
        off += align256(p.bytes);
    }
    // Solver.cu:13-104: uniform initial u, v, w, p, temp; everything else zero.
    e = hipMemsetAsync(h->arena, 0, total, h->stream);
    std::vector<float> host(N);
    // u, v, w also into their smoothers' partner buffers, which stay equal to them
    float *initDst[8] = {P.u, P.v, P.w, P.p, P.temp, P.alt[0], P.alt[1], P.alt[2]};
    const int initSrc[8] = {0, 1, 2, 3, 4, 0, 1, 2};
    for (int q = 0; q < 8 && e == hipSuccess; ++q) {
        std::fill(host.begin(), host.end(), (float)cfg->init[initSrc[q]]);
        e = hipMemcpyAsync(initDst[q], host.data(), N * 4, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) {
        std::string m = std::string("device initialisation: ") + hipGetErrorString(e);
        ifl_destroy(h);
        return fail(IFL_ERR_DEVICE, m);
    }
    *out = h;
    return IFL_OK;
}

int ifl_destroy(ifl_solver *h) {
    if (!h) return IFL_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->adj) {
        if (h->adj->graph) (void)hipGraphExecDestroy(h->adj->graph);
        if (h->adj->arena) (void)hipFree(h->adj->arena);
        delete h->adj;
    }
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    for (int f = 0; f < 6; ++f)
        if (h->prof[f]) (void)hipFree(h->prof[f]);
    if (h->arena) (void)hipFree(h->arena);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return IFL_OK;
}

int ifl_set_face_profile(ifl_solver *h, int face, const double *vals, size_t n) {
    if (!h || face < 0 || face > 5) return fail(IFL_ERR_ARG, "bad handle or face");
    const ifl_config &c = h->cfg;
    size_t need = face < 2 ? (size_t)c.ny * c.nz : face < 4 ? (size_t)c.nx * c.nz : (size_t)c.nx * c.ny;
    IFL_HIP(hipStreamSynchronize(h->stream));
    if (!vals) {
        if (h->prof[face]) IFL_HIP(hipFree(h->prof[face]));
        h->prof[face] = nullptr;
    } else {
        if (n != need) return fail(IFL_ERR_ARG, "profile length must equal the number of cells on the face");
        if (!h->prof[face]) IFL_HIP(hipMalloc((void **)&h->prof[face], need * 4));
        std::vector<float> tmp(n);
        for (size_t q = 0; q < n; ++q) tmp[q] = (float)vals[q];
        IFL_HIP(hipMemcpy(h->prof[face], tmp.data(), n * 4, hipMemcpyHostToDevice));
    }
    h->P.prof[face] = h->prof[face];
    h->graphDirty = true;
    return IFL_OK;
}

size_t ifl_field_size(const ifl_solver *h, int field) { return h ? fieldSize(h->cfg, field) : 0; }

int ifl_get_field(ifl_solver *h, int field, float *dst, size_t n) {
    if (!h || !dst) return fail(IFL_ERR_ARG, "null argument");
    float *src = fieldPtr(h, field);
    if (!src) return fail(IFL_ERR_ARG, field == IFL_FIELD_TCOEF || field == IFL_FIELD_BT ? "field needs energy = 1" : "bad field id");
    if (n != fieldSize(h->cfg, field)) return fail(IFL_ERR_ARG, "size mismatch");
    IFL_HIP(hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToHost, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    return IFL_OK;
}

int ifl_set_field(ifl_solver *h, int field, const float *src, size_t n) {
    if (!h || !src) return fail(IFL_ERR_ARG, "null argument");
    float *dst = fieldPtr(h, field);
    if (!dst) return fail(IFL_ERR_ARG, field == IFL_FIELD_TCOEF || field == IFL_FIELD_BT ? "field needs energy = 1" : "bad field id");
    if (n != fieldSize(h->cfg, field)) return fail(IFL_ERR_ARG, "size mismatch");
    IFL_HIP(hipMemcpyAsync(dst, src, n * 4, hipMemcpyHostToDevice, h->stream));
    if (field == IFL_FIELD_U || field == IFL_FIELD_V || field == IFL_FIELD_W)   // and the partner buffer
        IFL_HIP(hipMemcpyAsync(h->P.alt[field - IFL_FIELD_U], src, n * 4, hipMemcpyHostToDevice, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    return IFL_OK;
}

int ifl_solve(ifl_solver *h, int max_iter, ifl_report *report, ifl_iter_record *history, size_t hist_cap) {
    if (!h) return fail(IFL_ERR_ARG, "null handle");
    if (max_iter <= 0) max_iter = h->cfg.num_outer_iter;
    if (report) std::memset(report, 0, sizeof(*report));
    int st = buildGraph(h);
    if (st != IFL_OK) return st;
    st = armCase(h, h->stream, 0);
    if (st != IFL_OK) return st;
    SolveTally t;
    t.it0 = h->iter;
    t.maxIter = max_iter;
    t.history = history;
    t.histCap = hist_cap;
    const long long it0 = h->iter;
    int chunk = h->cfg.check_every > 0 ? std::min(h->cfg.check_every, kHistCap) : 64;
    std::vector<ifl_iter_record> ring;
    IFL_HIP(hipEventRecord(h->ev0, h->stream));
    if (h->iter == 0) ifl::launch_init_faces(planOf(h).L);           // Solver.cu:162-164
    long long launched = 0;
    ifl::CaseState s{};
    while (launched < max_iter) {
        int n = (int)std::min<long long>(chunk, max_iter - launched);
        for (int q = 0; q < n; ++q) IFL_HIP(hipGraphLaunch(h->graph, h->stream));
        launched += n;
        st = readState(h, s);
        if (st != IFL_OK) return st;
        const long long a = h->iter, b = s.iter;
        if (b > a) {
            ring.resize((size_t)(b - a));
            if ((st = copyHistory(h, h->stream, a, b, ring.data())) != IFL_OK) return st;
            IFL_HIP(hipStreamSynchronize(h->stream));
            tally(h, t, ring.data(), a, b, s.done != 0, h->cfg.log_every);
        }
        h->iter = b;
        if (s.done) break;
    }
    const double sweepsTotal = t.sweeps;
    IFL_HIP(hipEventRecord(h->ev1, h->stream));
    IFL_HIP(hipEventSynchronize(h->ev1));
    float ms = 0;
    IFL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    if (report) {
        report->iterations = (int)(h->iter - it0);
        report->converged = s.done == 1;
        report->status = s.done == 2 ? IFL_ERR_DIVERGED : IFL_OK;
        report->elapsed_ms = ms;
        report->final_scaled = s.lastScaled;
        report->sweeps_total = sweepsTotal;
    }
    if (s.done == 2) return fail(IFL_ERR_DIVERGED, "non-finite residual: the solve diverged");
    return IFL_OK;
}

int ifl_run_stage(ifl_solver *h, int stage, double *out, size_t n_out) {
    if (!h) return fail(IFL_ERR_ARG, "null handle");
    double tmp[8] = {0};
    int st = setDone(h, 0);
    if (st != IFL_OK) return st;
    const ifl::CaseParams &P = h->P;
    const Plan pl = planOf(h);
    const ifl::Lc &L = pl.L;
    ifl::CaseState s0{}, s{};
    st = readState(h, s0);
    if (st != IFL_OK) return st;
    switch (stage) {
        case IFL_STAGE_INIT_FACES: ifl::launch_init_faces(L); break;
        case IFL_STAGE_FACE_BC: ifl::launch_face_bc(L); break;
        case IFL_STAGE_MOM_ASSEMBLE:
            ifl::launch_mom_assemble(L);
            ifl::launch_reduce_cells(L, ifl::RED_MOM);
            break;
        case IFL_STAGE_MOM_RESIDUAL: break;
        case IFL_STAGE_SMOOTH_U:
        case IFL_STAGE_SMOOTH_V:
        case IFL_STAGE_SMOOTH_W: {
            const int slot = stage - IFL_STAGE_SMOOTH_U;
            if (slot == 2 && P.dim != 3) return fail(IFL_ERR_ARG, "SMOOTH_W needs dim 3");
            enqueueSmoothers(pl, &slot, 1);
            break;
        }
        case IFL_STAGE_SMOOTH_T:
            if (!P.energy) return fail(IFL_ERR_ARG, "SMOOTH_T needs energy = 1");
            enqueueTemperatureSmoother(pl);
            break;
        case IFL_STAGE_RHIE_CHOW: ifl::launch_rhie_chow(L); break;
        case IFL_STAGE_PCORR_ASSEMBLE:
            ifl::launch_pcorr_assemble(L);
            ifl::launch_reduce_cells(L, ifl::RED_MASS);
            break;
        case IFL_STAGE_SMOOTH_P: enqueuePressureCorrection(pl); break;
        case IFL_STAGE_UPDATE:
            ifl::launch_update(L);
            ifl::launch_reduce_cells(L, ifl::RED_UPD);
            break;
        case IFL_STAGE_OUTER:
            if (h->iter == 0) ifl::launch_init_faces(L);
            enqueueOuter(pl);
            break;
        default: return fail(IFL_ERR_ARG, "bad stage");
    }
    IFL_HIP(hipGetLastError());
    st = readState(h, s);
    if (st != IFL_OK) return st;
    switch (stage) {
        case IFL_STAGE_MOM_RESIDUAL:
            for (int q = 0; q < 4; ++q) tmp[q] = std::sqrt(s.resSq[q]);
            break;
        case IFL_STAGE_SMOOTH_T:
            tmp[0] = s.sm[ifl::kSlotT].sweeps;
            tmp[1] = s.sm[ifl::kSlotT].lastSum;
            break;
        case IFL_STAGE_SMOOTH_U:
        case IFL_STAGE_SMOOTH_V:
        case IFL_STAGE_SMOOTH_W: {
            int slot = stage - IFL_STAGE_SMOOTH_U;
            tmp[0] = s.sm[slot].sweeps;
            tmp[1] = s.sm[slot].lastSum;
            break;
        }
        case IFL_STAGE_PCORR_ASSEMBLE: tmp[0] = s.mass; break;
        case IFL_STAGE_SMOOTH_P:
            tmp[0] = s.sweepsIter[3] - s0.sweepsIter[3];
            tmp[1] = h->cfg.p_refine > 0 ? s.sm[4].lastSum : s.sm[3].lastSum;
            break;
        case IFL_STAGE_UPDATE:
            for (int q = 0; q < 7; ++q) tmp[q] = s.upd[q];
            break;
        case IFL_STAGE_OUTER:
            tmp[0] = s.lastConverged;
            tmp[1] = s.lastScaled;
            h->iter = s.iter;
            break;
        default: break;
    }
    if (out)
        for (size_t q = 0; q < n_out && q < 8; ++q) out[q] = tmp[q];
    return IFL_OK;
}

int ifl_time_kernel(ifl_solver *h, int kernel, int reps, double *avg_us, int *launches_per_rep) {
    if (!h || reps <= 0 || !avg_us) return fail(IFL_ERR_ARG, "bad argument");
    int st = setDone(h, 0);
    if (st != IFL_OK) return st;
    const Plan pl = planOf(h);
    const ifl::Lc &L = pl.L;
    const int per = 1;
    auto once = [&]() {
        switch (kernel) {
            case IFL_KERNEL_SWEEP: {
                // one launch of the p' smoother as the outer iteration runs it (first launch of a call)
                const int K = sweepBlockOf(h->cfg, pl.dim, 3);
                const ifl::SweepSel e{3, 3, K, 0};
                if (h->cfg.smoother == IFL_SMOOTHER_JACOBI) ifl::launch_sweep_jacobi(L, e, 0);
                else if (pl.dim == 2) ifl::launch_rb_sweep2d(L, ifl::COEF_PCORR, e, e, 1, 0, K);
                else ifl::launch_rb_sweep3d(L, ifl::COEF_PCORR, &e, 1, 0);
                break;
            }
            case IFL_KERNEL_MOM_SWEEP: {
                // one launch of the momentum smoother as the outer iteration runs it
                ifl::SweepSel se[3];
                const int K = sweepBlockOf(h->cfg, pl.dim, 0);
                for (int q = 0; q < 3; ++q) se[q] = ifl::SweepSel{q, q, K, 0};
                if (h->cfg.smoother == IFL_SMOOTHER_JACOBI) ifl::launch_sweep_jacobi(L, se[0], 0);
                else if (pl.dim == 2) ifl::launch_rb_sweep2d(L, ifl::COEF_MOMENTUM, se[0], se[1], 2, 0, K);
                else ifl::launch_rb_sweep3d(L, ifl::COEF_MOMENTUM, se, 3, 0);
                break;
            }
            case IFL_KERNEL_TEMP_SWEEP: {
                // one launch of the temperature smoother as the outer iteration runs it
                const int K = sweepBlockOf(h->cfg, pl.dim, ifl::kSlotT);
                const ifl::SweepSel e{ifl::kSlotT, 4, K, 0};
                if (h->cfg.smoother == IFL_SMOOTHER_JACOBI) ifl::launch_sweep_jacobi(L, e, 0);
                else if (pl.dim == 2) ifl::launch_rb_sweep2d(L, ifl::COEF_TEMP, e, e, 1, 0, K);
                else ifl::launch_rb_sweep3d(L, ifl::COEF_TEMP, &e, 1, 0);
                break;
            }
            case IFL_KERNEL_MOM_ASSEMBLE: ifl::launch_mom_assemble(L); break;
            case IFL_KERNEL_RHIE_CHOW: ifl::launch_rhie_chow(L); break;
            case IFL_KERNEL_PCORR_ASSEMBLE: ifl::launch_pcorr_assemble(L); break;
            case IFL_KERNEL_UPDATE: ifl::launch_update(L); break;
            default: break;
        }
    };
    if (kernel < 0 || kernel >= IFL_KERNEL_COUNT) return fail(IFL_ERR_ARG, "bad kernel id");
    if (kernel == IFL_KERNEL_TEMP_SWEEP && !h->P.energy) return fail(IFL_ERR_ARG, "the temperature sweep needs energy = 1");
    once();   // warm-up
    IFL_HIP(hipGetLastError());
    // one event pair per launch: the per-dispatch device duration, as a kernel trace reports it
    double total = 0.0;
    for (int r = 0; r < reps; ++r) {
        IFL_HIP(hipEventRecord(h->ev0, h->stream));
        once();
        IFL_HIP(hipEventRecord(h->ev1, h->stream));
        IFL_HIP(hipEventSynchronize(h->ev1));
        float ms = 0;
        IFL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        total += ms;
    }
    *avg_us = 1000.0 * total / ((double)reps * per);
    if (launches_per_rep) *launches_per_rep = per;
    return IFL_OK;
}

int ifl_write_vtk(ifl_solver *h, const char *path) {
    if (!h || !path) return fail(IFL_ERR_ARG, "null argument");
    const ifl_config &c = h->cfg;
    const size_t N = (size_t)h->P.N;
    std::vector<float> u(N), v(N), w(N), p(N), temp(N);
    int st;
    if ((st = ifl_get_field(h, IFL_FIELD_U, u.data(), N)) || (st = ifl_get_field(h, IFL_FIELD_V, v.data(), N))
        || (st = ifl_get_field(h, IFL_FIELD_W, w.data(), N)) || (st = ifl_get_field(h, IFL_FIELD_P, p.data(), N))
        || (st = ifl_get_field(h, IFL_FIELD_TEMP, temp.data(), N)))
        return st;
    std::filesystem::path fp(path);
    std::error_code ec;
    if (fp.has_parent_path() && !std::filesystem::exists(fp.parent_path()))
        std::filesystem::create_directories(fp.parent_path(), ec);
    std::ofstream file(fp);
    if (!file.is_open()) return fail(IFL_ERR_IO, std::string("Failed to open the file: ") + path);
    // Solver.cu:317-364 (legacy VTK ASCII, float coordinates, default stream precision)
    const float xmin = (float)c.xmin, ymin = (float)c.ymin, zmin = (float)c.zmin;
    const float dx = h->P.dx, dy = h->P.dy, dz = h->P.dz;
    file << "# vtk DataFile Version 3.0" << std::endl;
    file << "flash 3d grid and solution" << std::endl;
    file << "ASCII" << std::endl;
    file << "DATASET RECTILINEAR_GRID" << std::endl;
    file << "DIMENSIONS " << c.nx + 1 << " " << c.ny + 1 << " " << c.nz + 1 << std::endl;
    file << "X_COORDINATES " << c.nx + 1 << " float" << std::endl;
    for (int i = 0; i <= c.nx; ++i) file << xmin + (float)i * dx << " ";
    file << std::endl;
    file << "Y_COORDINATES " << c.ny + 1 << " float" << std::endl;
    for (int i = 0; i <= c.ny; ++i) file << ymin + (float)i * dy << " ";
    file << std::endl;
    file << "Z_COORDINATES " << c.nz + 1 << " float" << std::endl;
    for (int i = 0; i <= c.nz; ++i) file << zmin + (float)i * dz << " ";
    file << std::endl;
    const int ncell = c.nx * c.ny * c.nz;
    file << "CELL_DATA " << ncell << std::endl;
    file << "FIELD FieldData 3" << std::endl;
    file << "temperature 1 " << ncell << " float" << std::endl;
    for (const float &t : temp) file << t << " ";
    file << std::endl;
    file << "velocity 3 " << ncell << " float" << std::endl;
    for (int idx = 0; idx < ncell; ++idx) file << u[idx] << " " << v[idx] << " " << w[idx] << " ";
    file << std::endl;
    file << "pressure 1 " << ncell << " float" << std::endl;
    for (const float &x : p) file << x << " ";
    file.close();
    if (!file) return fail(IFL_ERR_IO, std::string("write failed: ") + path);
    return IFL_OK;
}

}  // extern "C"

// ---------------- binary state files ----------------
namespace {

constexpr int kStateFields[8] = {IFL_FIELD_U, IFL_FIELD_V, IFL_FIELD_W, IFL_FIELD_P,
                                 IFL_FIELD_TEMP, IFL_FIELD_UF, IFL_FIELD_VF, IFL_FIELD_WF};
static_assert(sizeof(ifl_state_header) == 160, "ifl_state_header layout");

uint64_t fnv1a(const unsigned char *p, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

size_t statePayloadFloats(const ifl_config &c) {
    size_t n = 0;
    for (int f : kStateFields) n += fieldSize(c, f);
    return n;
}

}  // namespace

extern "C" {

int ifl_save_state(ifl_solver *h, const char *path) {
    if (!h || !path) return fail(IFL_ERR_ARG, "null argument");
    const ifl_config &c = h->cfg;
    std::vector<float> data(statePayloadFloats(c));
    size_t off = 0;
    for (int f : kStateFields) {
        const size_t n = fieldSize(c, f);
        IFL_HIP(hipMemcpyAsync(data.data() + off, fieldPtr(h, f), n * 4, hipMemcpyDeviceToHost, h->stream));
        off += n;
    }
    ifl::CaseState s{};
    int st = readState(h, s);   // synchronises the stream (the field copies above too)
    if (st != IFL_OK) return st;
    ifl_state_header hd{};
    std::memcpy(hd.magic, IFL_STATE_MAGIC, 8);
    hd.version = IFL_STATE_VERSION;
    hd.dim = c.dim; hd.nx = c.nx; hd.ny = c.ny; hd.nz = c.nz;
    hd.last_converged = s.lastConverged;
    hd.iter = s.iter;
    const double nm[7] = {s.pNormMax, s.uNormMax, s.vNormMax, s.wNormMax, s.ufNormMax, s.vfNormMax, s.wfNormMax};
    for (int q = 0; q < 7; ++q) hd.norm_max[q] = nm[q];
    for (int q = 0; q < 5; ++q) hd.res_max[q] = s.resMax[q];
    hd.last_scaled = s.lastScaled;
    hd.payload_bytes = data.size() * 4;
    hd.checksum = fnv1a(reinterpret_cast<const unsigned char *>(data.data()), data.size() * 4);
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(IFL_ERR_IO, std::string("cannot open for writing: ") + path);
    const bool ok = std::fwrite(&hd, sizeof(hd), 1, fp) == 1
                    && std::fwrite(data.data(), 4, data.size(), fp) == data.size();
    const bool closed = std::fclose(fp) == 0;
    if (!ok || !closed) return fail(IFL_ERR_IO, std::string("write failed: ") + path);
    return IFL_OK;
}

int ifl_load_state(ifl_solver *h, const char *path) {
    if (!h || !path) return fail(IFL_ERR_ARG, "null argument");
    const ifl_config &c = h->cfg;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(IFL_ERR_IO, std::string("cannot open: ") + path);
    ifl_state_header hd{};
    std::vector<float> data(statePayloadFloats(c));
    const bool gotHeader = std::fread(&hd, sizeof(hd), 1, fp) == 1;
    int st = IFL_OK;
    if (!gotHeader || std::memcmp(hd.magic, IFL_STATE_MAGIC, 8) != 0)
        st = fail(IFL_ERR_IO, std::string("not a state file: ") + path);
    else if (hd.version != IFL_STATE_VERSION)
        st = fail(IFL_ERR_IO, "state file version " + std::to_string(hd.version) + ", expected "
                                  + std::to_string(IFL_STATE_VERSION));
    else if (hd.dim != c.dim || hd.nx != c.nx || hd.ny != c.ny || hd.nz != c.nz)
        st = fail(IFL_ERR_ARG, "state file grid " + std::to_string(hd.dim) + "D " + std::to_string(hd.nx) + "x"
                                   + std::to_string(hd.ny) + "x" + std::to_string(hd.nz) + " does not match the handle's");
    else if (hd.payload_bytes != data.size() * 4 || std::fread(data.data(), 4, data.size(), fp) != data.size()
             || std::fgetc(fp) != EOF)
        st = fail(IFL_ERR_IO, std::string("truncated or oversized state file: ") + path);
    else if (fnv1a(reinterpret_cast<const unsigned char *>(data.data()), data.size() * 4) != hd.checksum)
        st = fail(IFL_ERR_IO, std::string("state file checksum mismatch: ") + path);
    std::fclose(fp);
    if (st != IFL_OK) return st;
    ifl::CaseState s{};
    if ((st = readState(h, s)) != IFL_OK) return st;
    s.iter = hd.iter;
    s.lastConverged = hd.last_converged;
    double *nm[7] = {&s.pNormMax, &s.uNormMax, &s.vNormMax, &s.wNormMax, &s.ufNormMax, &s.vfNormMax, &s.wfNormMax};
    for (int q = 0; q < 7; ++q) *nm[q] = hd.norm_max[q];
    for (int q = 0; q < 5; ++q) s.resMax[q] = hd.res_max[q];
    s.lastScaled = hd.last_scaled;
    s.done = 0;
    size_t off = 0;
    for (int f : kStateFields) {
        const size_t n = fieldSize(c, f);
        IFL_HIP(hipMemcpyAsync(fieldPtr(h, f), data.data() + off, n * 4, hipMemcpyHostToDevice, h->stream));
        if (f == IFL_FIELD_U || f == IFL_FIELD_V || f == IFL_FIELD_W)   // the smoothers' partner buffers
            IFL_HIP(hipMemcpyAsync(h->P.alt[f - IFL_FIELD_U], data.data() + off, n * 4, hipMemcpyHostToDevice,
                                   h->stream));
        off += n;
    }
    IFL_HIP(hipMemcpyAsync(h->P.st, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    h->iter = hd.iter;
    return IFL_OK;
}

}  // extern "C"

// ---------------- autotuner ----------------
namespace {

// A copy of the whole device arena (fields, matrices, monitor, history) and the iteration count.
struct Snapshot {
    char *buf = nullptr;
    size_t bytes = 0;
    long long iter = 0;
    ~Snapshot() {
        if (buf) (void)hipFree(buf);
    }
};

int snapTake(ifl_solver *h, Snapshot &s) {
    s.bytes = h->arenaBytes;
    s.iter = h->iter;
    IFL_HIP(hipMalloc((void **)&s.buf, s.bytes));
    IFL_HIP(hipMemcpyAsync(s.buf, h->arena, s.bytes, hipMemcpyDeviceToDevice, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    return IFL_OK;
}

int snapRestore(ifl_solver *h, const Snapshot &s) {
    IFL_HIP(hipMemcpyAsync(h->arena, s.buf, s.bytes, hipMemcpyDeviceToDevice, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    h->iter = s.iter;
    return IFL_OK;
}

struct Settings {
    int tile, K, nIter[4], stopRule, Km;
};

void applySettings(ifl_solver *h, const Settings &x) {
    h->cfg.tile = x.tile;
    if (h->P.dim == 2) h->tile = x.tile;
    h->cfg.sweep_block = x.K;
    h->cfg.sweep_block_mom = x.Km;
    for (int q = 0; q < 4; ++q) h->cfg.n_iter[q] = x.nIter[q];
    h->cfg.stop_rule = x.stopRule;
    h->P.stopRule = x.stopRule;
    h->graphDirty = true;
}

// n outer iterations from the current state: device time, final scaled residual, divergence.
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
    h->iter = s.iter;
    scaled = s.lastScaled;
    diverged = s.done == 2 || !std::isfinite(scaled);
    return IFL_OK;
}

}  // namespace

extern "C" int ifl_autotune(ifl_solver *h, int what, int trial_iters, ifl_tune_result *out) {
    if (!h || !out || (what & ~(IFL_TUNE_LAUNCH | IFL_TUNE_SWEEPS))) return fail(IFL_ERR_ARG, "bad argument");
    std::memset(out, 0, sizeof(*out));
    const int trial = trial_iters > 0 ? trial_iters : 300;
    const ifl_config c0 = h->cfg;
    Settings best{h->P.dim == 2 ? h->tile : 0, std::max(1, c0.sweep_block), {}, IFL_STOP_NONE, c0.sweep_block_mom};
    for (int q = 0; q < 4; ++q) best.nIter[q] = c0.n_iter[q];
    int st;
    double ms = 0, sc = 0;
    bool div = false;
    applySettings(h, best);
    if (h->iter < 20 && (st = trialRun(h, (int)(20 - h->iter), ms, sc, div)) != IFL_OK) return st;
    Snapshot snap;
    if ((st = snapTake(h, snap)) != IFL_OK) return st;
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
                applySettings(h, x);
                if ((st = trialRun(h, trial, ms, sc, div)) != IFL_OK) return st;
                if ((st = snapRestore(h, snap)) != IFL_OK) return st;
                out->trials++;
                if (div || !(sc > 0.0) || !(s0.lastScaled > 0.0)) continue;
                const double rate = std::log10(s0.lastScaled / sc) / (ms / 1000.0);
                if (rate > bestRate) {
                    bestRate = rate;
                    chosen = x;
                }
            }
        best = chosen;
        if (bestRate == -1e300) bestRate = 0.0;
    }
    if (!(what & IFL_TUNE_LAUNCH) || bestUs == 0.0 || (what & IFL_TUNE_SWEEPS))
        if ((st = timeIter(best, bestUs)) != IFL_OK) return st;
    best.stopRule = c0.stop_rule;
    applySettings(h, best);
    out->tile = best.tile;
    out->sweep_block = best.K;
    out->sweep_block_mom = best.Km;
    for (int q = 0; q < 4; ++q) out->n_iter[q] = best.nIter[q];
    out->us_per_iter = bestUs;
    out->decades_per_s = bestRate;
    return IFL_OK;
}

// ---------------- discrete adjoint ----------------
namespace {

int adjAlloc(ifl_solver *h) {
    if (h->adj) return IFL_OK;
    AdjHost *a = new (std::nothrow) AdjHost();
    if (!a) return fail(IFL_ERR_NOMEM, "out of host memory");
    h->adj = a;
    const ifl::CaseParams &P = h->P;
    const size_t N = (size_t)P.N, nx = P.nx, ny = P.ny, nz = P.nz;
    const size_t nblk = (size_t)ifl::cell_blocks(P);
    ifl::AdjBufs &B = a->B;
    struct Part { void **dst; size_t bytes; };
    std::vector<Part> parts;
    for (int d = 0; d < 3; ++d) {
        parts.push_back({(void **)&B.lu[d], N * 4});
        parts.push_back({(void **)&B.us[d], N * 4});
        parts.push_back({(void **)&B.usAlt[d], N * 4});
        parts.push_back({(void **)&B.rb[d], N * 4});
    }
    parts.push_back({(void **)&B.lf[0], (nx + 1) * ny * nz * 4});
    parts.push_back({(void **)&B.lf[1], nx * (ny + 1) * nz * 4});
    parts.push_back({(void **)&B.lf[2], nx * ny * (nz + 1) * 4});
    parts.push_back({(void **)&B.lp, N * 4});
    parts.push_back({(void **)&B.pcb, N * 4});
    parts.push_back({(void **)&B.pcbAlt, N * 4});
    parts.push_back({(void **)&B.bpc, N * 4});
    parts.push_back({(void **)&B.acb, N * 4});
    parts.push_back({(void **)&B.pacc, N * 4});
    parts.push_back({(void **)&B.abar, N * 7 * 4});
    for (int q = 0; q < 4; ++q) parts.push_back({(void **)&a->gw[q], N * 4});
    const size_t faceCells[6] = {ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny};
    for (int f = 0; f < 6; ++f) parts.push_back({(void **)&B.gprof[f], faceCells[f] * 8});
    parts.push_back({(void **)&B.part, nblk * 8 * 8});
    parts.push_back({(void **)&B.st, sizeof(ifl::AdjState)});
    size_t total = 0;
    for (auto &p : parts) total += align256(p.bytes);
    IFL_HIP(hipMalloc((void **)&a->arena, total));
    size_t off = 0;
    for (auto &p : parts) {
        *p.dst = a->arena + off;
        off += align256(p.bytes);
    }
    IFL_HIP(hipMemsetAsync(a->arena, 0, total, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    B.nblk = (int)nblk;
    return IFL_OK;
}

// Tile of the 2D transposed sweeps: the configured one, else 64x8 with two sweeps per launch
// (the fastest on a 1024x256 channel: 0.199 ms per adjoint iteration against 0.212-0.237 ms for
// the other shapes), or 32x8 with four when the grid has fewer than 256 such tiles.
int adjShape(const ifl::CaseParams &P, int tile) {
    if (tile > 0) return tile - 1;
    const long tiles = (long)((P.nx + 63) / 64) * ((P.ny + 7) / 8);
    return tiles >= 256 ? 2 : 3;
}

// n transposed sweeps of one equation from *x (the result ends in *x; *alt is its partner).
void enqueueAdjSweeps(const ifl::Lc &L, const ifl::AdjBufs &B, int tile, float *&x, float *&alt, float *bb,
                      const float *coef, float rlx, int n) {
    const ifl::CaseParams &P = *L.P;
    if (P.dim == 2) {
        const int shape = adjShape(P, tile), K = ifl::adj_tile_sweeps(shape);
        for (int s = 0; s < n; s += K) {
            ifl::launch_adj_tile2d(L, shape, x, alt, bb, coef, rlx, std::min(K, n - s));
            std::swap(x, alt);
        }
        return;
    }
    for (int s = 0; s < n; ++s) {
        ifl::launch_adj_sweep_t(L, B, x, alt, bb, coef, rlx, 1);
        ifl::launch_adj_sweep_t(L, B, alt, x, bb, coef, rlx, 0);
    }
}

// One adjoint iteration (reverse stage order of one outer iteration).  B is a copy: the sweeps
// swap its ping-pong pointers, and every iteration starts from the same assignment.
void enqueueAdjoint(ifl_solver *h, int tile) {
    const Plan pl = planOf(h);
    const ifl::Lc &L = pl.L;
    ifl::AdjBufs B = h->adj->B;
    const ifl::CaseParams &P = h->P;
    ifl::launch_adj_pcorr(L, B);
    enqueueAdjSweeps(L, B, tile, B.pcb, B.pcbAlt, B.bpc, P.pcoef, P.relax[3], h->cfg.n_iter[3]);
    ifl::launch_adj_rc(L, B);
    for (int d = 0; d < P.dim; ++d)
        enqueueAdjSweeps(L, B, tile, B.us[d], B.usAlt[d], B.rb[d], P.mcoef, P.relax[d], h->cfg.n_iter[d]);
    ifl::launch_adj_coef(L, B);
    ifl::launch_adj_face(L, B);
    ifl::launch_adj_reduce(L, B);
}

}  // namespace

extern "C" void ifl_adjoint_opts_default(ifl_adjoint_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->objective = IFL_OBJ_PRESSURE_DROP;
    o->max_iter = 200000;
    o->tol = 1e-7;
    o->check_every = 64;
}

extern "C" int ifl_adjoint(ifl_solver *h, const ifl_adjoint_opts *o, int face, double *dJ_dprofile, size_t n_profile,
                           double *dJ_dmu, ifl_adjoint_report *report) {
    if (!h || !o) return fail(IFL_ERR_ARG, "null argument");
    if (o->objective != IFL_OBJ_PRESSURE_DROP && o->objective != IFL_OBJ_LINEAR) return fail(IFL_ERR_ARG, "bad objective");
    if (o->tile < 0 || o->tile > ifl::kAdjShapes) return fail(IFL_ERR_ARG, "adjoint tile must be 0..6");
    const ifl_config &c = h->cfg;
    const size_t N = (size_t)h->P.N;
    if (dJ_dprofile) {
        if (face < 0 || face >= 2 * c.dim || c.bc_type[face] != IFL_INLET)
            return fail(IFL_ERR_ARG, "the profile gradient needs an inlet face");
        const size_t need = face < 2 ? (size_t)c.ny * c.nz : face < 4 ? (size_t)c.nx * c.nz : (size_t)c.nx * c.ny;
        if (n_profile != need) return fail(IFL_ERR_ARG, "profile gradient length must equal the cells on the face");
    }
    if (report) std::memset(report, 0, sizeof(*report));
    int st = adjAlloc(h);
    if (st != IFL_OK) return st;
    AdjHost &A = *h->adj;
    ifl::AdjBufs &B = A.B;
    // the fixed point: face BCs, momentum and p' matrices of the current state
    if ((st = armCase(h, h->stream, 0)) != IFL_OK) return st;
    const Plan pl = planOf(h);
    ifl::launch_face_bc(pl.L);
    ifl::launch_mom_assemble(pl.L);
    ifl::launch_pcorr_assemble(pl.L);
    IFL_HIP(hipGetLastError());
    // objective weights
    std::vector<float> w(N);
    for (int q = 0; q < 4; ++q) {
        const float *src = nullptr;
        if (o->objective == IFL_OBJ_LINEAR) src = o->weights[q];
        else if (q == 3) {
            std::fill(w.begin(), w.end(), 0.0f);
            const float v = 1.0f / (float)((long long)c.ny * c.nz);
            for (int k = 0; k < c.nz; ++k)
                for (int j = 0; j < c.ny; ++j) {
                    w[(size_t)c.nx * (j + (size_t)c.ny * k)] += v;
                    w[(size_t)c.nx - 1 + (size_t)c.nx * (j + (size_t)c.ny * k)] -= v;
                }
            src = w.data();
        }
        if (src) {
            IFL_HIP(hipMemcpyAsync(A.gw[q], src, N * 4, hipMemcpyHostToDevice, h->stream));
            IFL_HIP(hipStreamSynchronize(h->stream));
        }
        if (q < 3) B.gu[q] = src ? A.gw[q] : nullptr;
        else B.gp = src ? A.gw[q] : nullptr;
    }
    if (!o->warm_start) {
        const size_t nl = (size_t)((char *)B.pcb - (char *)B.lf[0]);   // lf[0..2] and lp are contiguous
        IFL_HIP(hipMemsetAsync(B.lf[0], 0, nl, h->stream));
        for (int d = 0; d < 3; ++d) IFL_HIP(hipMemsetAsync(B.lu[d], 0, N * 4, h->stream));
        IFL_HIP(hipMemsetAsync(B.st, 0, sizeof(ifl::AdjState), h->stream));
        IFL_HIP(hipStreamSynchronize(h->stream));
    }
    // graph of one adjoint iteration (captured per call: sweep counts and weight pointers are
    // part of it)
    if (A.graph) {
        (void)hipGraphExecDestroy(A.graph);
        A.graph = nullptr;
    }
    hipGraph_t g = nullptr;
    IFL_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueueAdjoint(h, o->tile);
    IFL_HIP(hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(&A.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(IFL_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    const int maxIter = o->max_iter > 0 ? o->max_iter : 200000;
    const double tol = o->tol > 0 ? o->tol : (o->tol < 0 ? 0.0 : 1e-7);
    const int chunk = o->check_every > 0 ? o->check_every : 64;
    IFL_HIP(hipEventRecord(h->ev0, h->stream));
    ifl::AdjState as{};
    long long done = 0;
    bool conv = false;
    double prevMub = 0.0;
    while (done < maxIter) {
        const int n = (int)std::min<long long>(chunk, maxIter - done);
        for (int q = 0; q < n; ++q) IFL_HIP(hipGraphLaunch(A.graph, h->stream));
        done += n;
        IFL_HIP(hipMemcpyAsync(&as, B.st, sizeof(as), hipMemcpyDeviceToHost, h->stream));
        IFL_HIP(hipStreamSynchronize(h->stream));
        if (!std::isfinite(as.scaled) || !std::isfinite(as.mub)) return fail(IFL_ERR_DIVERGED, "non-finite adjoint");
        if (as.iters >= 5 && as.scaled < tol) {
            conv = true;
            break;
        }
        if (o->grad_tol > 0 && done > n && std::fabs(as.mub - prevMub) <= o->grad_tol * std::fabs(as.mub)) {
            conv = true;
            break;
        }
        prevMub = as.mub;
    }
    IFL_HIP(hipEventRecord(h->ev1, h->stream));
    IFL_HIP(hipEventSynchronize(h->ev1));
    float ms = 0;
    IFL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    if (dJ_dmu) *dJ_dmu = as.mub;
    if (dJ_dprofile) {
        IFL_HIP(hipMemcpyAsync(dJ_dprofile, B.gprof[face], n_profile * 8, hipMemcpyDeviceToHost, h->stream));
        IFL_HIP(hipStreamSynchronize(h->stream));
    }
    if (report) {
        // objective at the current state
        double J = 0.0;
        std::vector<float> f(N);
        for (int q = 0; q < 4; ++q) {
            const bool has = q < 3 ? B.gu[q] != nullptr : B.gp != nullptr;
            if (!has) continue;
            const int fid = q < 3 ? IFL_FIELD_U + q : IFL_FIELD_P;
            if ((st = ifl_get_field(h, fid, f.data(), N)) != IFL_OK) return st;
            std::vector<float> wq(N);
            IFL_HIP(hipMemcpy(wq.data(), A.gw[q], N * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; ++i) J += (double)wq[i] * (double)f[i];
        }
        report->iterations = (int)done;
        report->converged = conv ? 1 : 0;
        report->J = J;
        report->final_scaled = as.scaled;
        report->elapsed_ms = ms;
    }
    return IFL_OK;
}

extern "C" int ifl_adjoint_lambda(ifl_solver *h, int field, float *dst, size_t n) {
    if (!h || !dst || !h->adj) return fail(IFL_ERR_ARG, "no adjoint on this handle");
    const ifl::AdjBufs &B = h->adj->B;
    const float *src = nullptr;
    switch (field) {
        case IFL_FIELD_U: src = B.lu[0]; break;
        case IFL_FIELD_V: src = B.lu[1]; break;
        case IFL_FIELD_W: src = B.lu[2]; break;
        case IFL_FIELD_P: src = B.lp; break;
        case IFL_FIELD_UF: src = B.lf[0]; break;
        case IFL_FIELD_VF: src = B.lf[1]; break;
        case IFL_FIELD_WF: src = B.lf[2]; break;
        default: return fail(IFL_ERR_ARG, "bad adjoint field");
    }
    if (n != fieldSize(h->cfg, field)) return fail(IFL_ERR_ARG, "size mismatch");
    IFL_HIP(hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToHost, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    return IFL_OK;
}

// ---------------- batched solves ----------------
// The batch owns one solver handle per case (arena, monitor, history ring) and runs them all
// through one graph on its own stream.  A device table holds every case's CaseParams and, per
// workgroup kind, the prefix sums of the cases' workgroup counts; a kernel launch covers the sum
// and each workgroup finds its case by binary search.  When half the cases of the graph have
// stopped, the table and graph are rebuilt over the cases still running.
struct ifl_batch {
    std::vector<ifl_solver *> cases;
    ifl_config cfg{};                 // the settings that fix the launch sequence (from case 0)
    int tile = 0;
    hipStream_t stream = nullptr;
    char *table = nullptr;            // device: CaseParams[n], then start/dimx per kind
    size_t tableBytes = 0;
    std::vector<char> hostTable;
    hipGraphExec_t graph = nullptr;
    std::vector<int> inGraph;         // case indices the graph covers
    ifl::Lc L{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

int batchValidate(const ifl_config *c, int n) {
    for (int i = 0; i < n; ++i) {
        int st = validate(c[i]);
        if (st != IFL_OK) return fail(st, "case " + std::to_string(i) + ": " + g_err);
        const ifl_config &a = c[0], &b = c[i];
        bool same = a.dim == b.dim && a.smoother == b.smoother && a.p_refine == b.p_refine
                    && a.sweep_block == b.sweep_block && a.sweep_block_mom == b.sweep_block_mom
                    && a.tile == b.tile && a.energy == b.energy
                    && (!a.energy || a.n_iter_t == b.n_iter_t);
        for (int q = 0; q < 4; ++q) same = same && a.n_iter[q] == b.n_iter[q];
        if (!same)
            return fail(IFL_ERR_ARG, "case " + std::to_string(i)
                                         + ": dim, smoother, n_iter, p_refine, sweep_block(_mom), tile, energy and n_iter_t "
                                           "must agree across a batch");
    }
    return IFL_OK;
}

// Tile of a batch: the configured one, else the largest whose total tile count fills the GPU.
int batchTile(const ifl_batch *b) {
    if (b->cfg.dim != 2) return 0;
    if (b->cfg.tile) return b->cfg.tile;
    for (int t = ifl::TILE_64x16; t < ifl::TILE_32x4; ++t) {
        long long tot = 0;
        for (ifl_solver *h : b->cases) tot += ifl::tile_count(h->P.nx, h->P.ny, t);
        if (tot >= 768) return t;
    }
    return ifl::TILE_32x4;
}

// Upload the table for the cases `ids` and capture the outer-iteration graph over them.
int batchBuild(ifl_batch *b, const std::vector<int> &ids) {
    if (b->graph) {
        IFL_HIP(hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
    }
    const int n = (int)ids.size();
    const size_t pBytes = align256(sizeof(ifl::CaseParams) * n);
    const size_t kBytes = sizeof(int) * (2 * n + 1);
    // per-workgroup tables: count the workgroups of every kind first
    size_t blkBytes[ifl::MAP_KINDS];
    for (int k = 0; k < ifl::MAP_KINDS; ++k) {
        long long tot = 0;
        for (int c = 0; c < n; ++c) {
            const ifl::CaseParams &P = b->cases[ids[c]]->P;
            switch (k) {
                case ifl::MAP_CELL: tot += ifl::cell_blocks(P); break;
                case ifl::MAP_TILE: tot += P.dim == 2 ? ifl::tile_count(P.nx, P.ny, b->tile ? b->tile : ifl::TILE_64x16) : 0; break;
                case ifl::MAP_COPY: tot += ifl::copy_blocks(P); break;
                case ifl::MAP_ZMARCH: tot += P.dim == 3 ? ifl::zmarch_blocks(P, false) : 0; break;
                case ifl::MAP_ZMARCH_P: tot += P.dim == 3 ? ifl::zmarch_blocks(P, true) : 0; break;
                default: tot += 1; break;
            }
        }
        blkBytes[k] = align256(sizeof(int4) * (size_t)std::max(1LL, tot));
    }
    size_t need = pBytes + ifl::MAP_KINDS * align256(kBytes);
    for (int k = 0; k < ifl::MAP_KINDS; ++k) need += blkBytes[k];
    if (need > b->tableBytes) {
        if (b->table) IFL_HIP(hipFree(b->table));
        b->table = nullptr;
        b->tableBytes = 0;
        IFL_HIP(hipMalloc((void **)&b->table, need));
        b->tableBytes = need;
    }
    std::vector<char> &ht = b->hostTable;
    ht.assign(need, 0);
    ifl::CaseParams *hp = (ifl::CaseParams *)ht.data();
    ifl::Lc &L = b->L;
    L = ifl::Lc{};
    L.batch = true;
    L.tile = b->tile ? b->tile : ifl::TILE_64x16;
    L.s = b->stream;
    L.P = &b->cases[ids[0]]->P;
    L.M.ncase = n;
    L.M.P = (const ifl::CaseParams *)b->table;
    for (int k = 0; k < ifl::MAP_KINDS; ++k) {
        const size_t off = pBytes + k * align256(kBytes);
        int *start = (int *)(ht.data() + off), *dimx = start + n + 1;
        long long acc = 0;
        for (int c = 0; c < n; ++c) {
            const ifl::CaseParams &P = b->cases[ids[c]]->P;
            int cnt = 1, dx = 1;
            switch (k) {
                case ifl::MAP_CELL: cnt = ifl::cell_blocks(P); dx = ifl::cell_dimx(P); break;
                case ifl::MAP_TILE:
                    cnt = P.dim == 2 ? ifl::tile_count(P.nx, P.ny, L.tile) : 0;
                    dx = P.dim == 2 ? ifl::tile_dimx(P.nx, L.tile) : 1;
                    break;
                case ifl::MAP_COPY: cnt = ifl::copy_blocks(P); dx = cnt; break;
                case ifl::MAP_ZMARCH: cnt = P.dim == 3 ? ifl::zmarch_blocks(P, false) : 0; dx = ifl::zmarch_dimx(P, false); break;
                case ifl::MAP_ZMARCH_P: cnt = P.dim == 3 ? ifl::zmarch_blocks(P, true) : 0; dx = ifl::zmarch_dimx(P, true); break;
                default: break;
            }
            start[c] = (int)acc;
            dimx[c] = dx;
            acc += cnt;
        }
        if (acc >= (1LL << 31)) return fail(IFL_ERR_ARG, "batch too large for one launch");
        start[n] = (int)acc;
        L.total[k] = (int)acc;
        L.M.start[k] = (const int *)(b->table + off);
        L.M.dimx[k] = (const int *)(b->table + off) + n + 1;
    }
    size_t boff = pBytes + ifl::MAP_KINDS * align256(kBytes);
    for (int k = 0; k < ifl::MAP_KINDS; ++k) {
        const int *start = (const int *)(ht.data() + pBytes + k * align256(kBytes));
        const int *dimx = start + n + 1;
        int4 *e = (int4 *)(ht.data() + boff);
        for (int c = 0; c < n; ++c)
            for (int w = start[c]; w < start[c + 1]; ++w) e[w] = make_int4(c, w - start[c], start[c + 1] - start[c], dimx[c]);
        L.M.blk[k] = (const int4 *)(b->table + boff);
        boff += blkBytes[k];
    }
    for (int c = 0; c < n; ++c) hp[c] = b->cases[ids[c]]->P;
    IFL_HIP(hipMemcpyAsync(b->table, ht.data(), need, hipMemcpyHostToDevice, b->stream));
    IFL_HIP(hipStreamSynchronize(b->stream));
    Plan pl;
    pl.L = L;
    pl.cfg = &b->cfg;
    pl.dim = b->cfg.dim;
    const int st = captureGraph(b->stream, pl, &b->graph);
    if (st != IFL_OK) return st;
    b->inGraph = ids;
    return IFL_OK;
}

}  // namespace

extern "C" {

int ifl_batch_create(const ifl_config *cfgs, int n, ifl_batch **out) {
    if (!cfgs || !out || n <= 0) return fail(IFL_ERR_ARG, "null argument or empty batch");
    *out = nullptr;
    int st = batchValidate(cfgs, n);
    if (st != IFL_OK) return st;
    ifl_batch *b = new (std::nothrow) ifl_batch();
    if (!b) return fail(IFL_ERR_NOMEM, "out of host memory");
    b->cfg = cfgs[0];
    for (int i = 0; i < n; ++i) {
        ifl_solver *h = nullptr;
        st = ifl_create(&cfgs[i], &h);
        if (st != IFL_OK) {
            std::string m = "case " + std::to_string(i) + ": " + g_err;
            ifl_batch_destroy(b);
            return fail(st, m);
        }
        b->cases.push_back(h);
    }
    b->tile = batchTile(b);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        std::string m = std::string("batch stream: ") + hipGetErrorString(e);
        ifl_batch_destroy(b);
        return fail(IFL_ERR_DEVICE, m);
    }
    *out = b;
    return IFL_OK;
}

int ifl_batch_destroy(ifl_batch *b) {
    if (!b) return IFL_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->table) (void)hipFree(b->table);
    for (ifl_solver *h : b->cases) ifl_destroy(h);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return IFL_OK;
}

int ifl_batch_size(const ifl_batch *b) { return b ? (int)b->cases.size() : 0; }

int ifl_batch_case(ifl_batch *b, int i, ifl_solver **out) {
    if (!b || !out || i < 0 || i >= (int)b->cases.size()) return fail(IFL_ERR_ARG, "bad batch or case index");
    *out = b->cases[i];
    return IFL_OK;
}

int ifl_batch_solve(ifl_batch *b, int max_iter, ifl_report *reports) {
    if (!b) return fail(IFL_ERR_ARG, "null batch");
    const int n = (int)b->cases.size();
    int limit = max_iter;
    if (limit <= 0)
        for (ifl_solver *h : b->cases) limit = std::max(limit, h->cfg.num_outer_iter);
    if (reports) std::memset(reports, 0, sizeof(ifl_report) * n);
    int st;
    for (ifl_solver *h : b->cases) IFL_HIP(hipStreamSynchronize(h->stream));
    // every case runs at most its own limit (num_outer_iter, or max_iter when given)
    std::vector<long long> it0(n);
    std::vector<SolveTally> t(n);
    for (int i = 0; i < n; ++i) {
        ifl_solver *h = b->cases[i];
        const int own = max_iter > 0 ? max_iter : h->cfg.num_outer_iter;
        it0[i] = h->iter;
        t[i].it0 = h->iter;
        t[i].maxIter = own;
        if ((st = armCase(h, b->stream, h->iter + own)) != IFL_OK) return st;
    }
    std::vector<int> live(n);
    for (int i = 0; i < n; ++i) live[i] = i;
    if ((st = batchBuild(b, live)) != IFL_OK) return st;
    const int chunk = b->cfg.check_every > 0 ? std::min(b->cfg.check_every, kHistCap) : 64;
    IFL_HIP(hipEventRecord(b->ev0, b->stream));
    for (ifl_solver *h : b->cases)
        if (h->iter == 0) {                                           // Solver.cu:162-164
            Plan pl = planOf(h);
            pl.L.s = b->stream;
            ifl::launch_init_faces(pl.L);
        }
    std::vector<ifl::CaseState> s(n);
    std::vector<int> finalDone(n, 0);
    std::vector<double> finalScaled(n, 0.0);
    std::vector<ifl_iter_record> ring;
    long long launched = 0;
    while (launched < limit && !live.empty()) {
        const int k = (int)std::min<long long>(chunk, limit - launched);
        for (int q = 0; q < k; ++q) IFL_HIP(hipGraphLaunch(b->graph, b->stream));
        launched += k;
        for (int i : live)
            IFL_HIP(hipMemcpyAsync(&s[i], b->cases[i]->P.st, sizeof(ifl::CaseState), hipMemcpyDeviceToHost, b->stream));
        IFL_HIP(hipStreamSynchronize(b->stream));
        std::vector<int> still;
        for (int i : live) {
            ifl_solver *h = b->cases[i];
            const long long a = h->iter, e = s[i].iter;
            if (e > a) {
                ring.resize((size_t)(e - a));
                if ((st = copyHistory(h, b->stream, a, e, ring.data())) != IFL_OK) return st;
                IFL_HIP(hipStreamSynchronize(b->stream));
                tally(h, t[i], ring.data(), a, e, s[i].done != 0, 0);
            }
            h->iter = e;
            finalDone[i] = s[i].done;
            finalScaled[i] = s[i].lastScaled;
            if (!s[i].done) still.push_back(i);
        }
        live.swap(still);
        if (!live.empty() && launched < limit && 2 * live.size() <= b->inGraph.size())
            if ((st = batchBuild(b, live)) != IFL_OK) return st;
    }
    IFL_HIP(hipEventRecord(b->ev1, b->stream));
    IFL_HIP(hipEventSynchronize(b->ev1));
    float ms = 0;
    IFL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    bool diverged = false;
    for (int i = 0; i < n; ++i) {
        if ((st = armCase(b->cases[i], b->stream, 0)) != IFL_OK) return st;
        if (finalDone[i] == 3) finalDone[i] = 0;
        diverged = diverged || finalDone[i] == 2;
        if (reports) {
            ifl_report &r = reports[i];
            r.iterations = (int)(b->cases[i]->iter - it0[i]);
            r.converged = finalDone[i] == 1;
            r.status = finalDone[i] == 2 ? IFL_ERR_DIVERGED : IFL_OK;
            r.elapsed_ms = ms;
            r.final_scaled = finalScaled[i];
            r.sweeps_total = t[i].sweeps;
        }
    }
    if (diverged) return fail(IFL_ERR_DIVERGED, "non-finite residual: a case of the batch diverged");
    return IFL_OK;
}

}  // extern "C"
