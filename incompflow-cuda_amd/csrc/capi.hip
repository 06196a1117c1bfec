// capi.hip - host side of the C-ABI in include/incompflow.h (C++20 + HIP runtime): handles,
// the launch plan of one outer iteration, solves, field access, stages, kernel timing and VTK.
//
// A handle owns one device arena (all fields, coefficient planes, reduction partials, the
// CaseState monitor and a ring of per-iteration records) and one HIP stream.  One SIMPLE outer
// iteration (Solver.cu:160-290) is captured once into a hipGraph and replayed; the device-side
// stop flag makes every kernel of an already finished case return at once, so the host reads the
// monitor only every `check_every` iterations instead of after every sweep (Solver.cu:951-953)
// and every iteration (Solver.cu:207-224).  State files, the autotuner, the adjoint and batches
// live in capi_state.hip, capi_tune.hip, capi_adjoint.hip and capi_batch.hip.
#include <filesystem>
#include <fstream>

#include "capi_host.hpp"

namespace ifl_host {

thread_local std::string g_err;

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
    if (c.fail_if_not_converged != 0 && c.fail_if_not_converged != 1)
        return fail(IFL_ERR_ARG, "fail_if_not_converged must be 0 or 1");
    if (c.p_solver != IFL_PSOLVE_SWEEPS && c.p_solver != IFL_PSOLVE_MG)
        return fail(IFL_ERR_ARG, "p_solver must be IFL_PSOLVE_SWEEPS or IFL_PSOLVE_MG");
    if (c.mg_cycles < 0 || c.mg_cycles > 8) return fail(IFL_ERR_ARG, "mg_cycles must be 0..8");
    if (c.mg_sweeps < 0 || c.mg_sweeps > 3) return fail(IFL_ERR_ARG, "mg_sweeps must be 0..3");
    if (c.mg_coarse_sweeps < 0 || c.mg_coarse_sweeps > 200) return fail(IFL_ERR_ARG, "mg_coarse_sweeps must be 0..200");
    if (!(c.mg_over >= 0.0 && c.mg_over <= 2.0)) return fail(IFL_ERR_ARG, "mg_over must be in [0, 2] (0: 1.0)");
    if (c.p_solver == IFL_PSOLVE_MG) {
        if (c.p_refine != 0) return fail(IFL_ERR_ARG, "p_refine needs p_solver = IFL_PSOLVE_SWEEPS");
        int dims[ifl::kMgMaxLevels][2];
        if (c.dim == 2 && ifl::mg_bottom_level(ifl::mg_levels(c.nx, c.ny, dims), dims) < 0)
            return fail(IFL_ERR_ARG, "IFL_PSOLVE_MG: the grid is too elongated (its coarsest level has more than 5600 cells)");
    }
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
    // rows per thread of the cell kernels: one on grids up to 2^19 cells, whose kernels are latency
    // bound (a 512x128 iteration 0.109 -> 0.100 ms, 1024x256 0.151 -> 0.142), four above (1024^2 and
    // 256^3 measured 1-2 % faster with four)
    P.cellRows = P.N <= (1 << 19) ? 1 : ifl::kRows;
    P.zcM = ifl::zmarch_chunk(P, false);
    P.zcP = ifl::zmarch_chunk(P, true);
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


// Sweeps per launch of a red-black smoother call (1 for Jacobi).  2D: sweep_block for p' and its
// refinement, sweep_block_mom for the momentum and temperature equations (0: the smaller of
// sweep_block and 2).  3D: sweep_block, at most 2, for p' and its refinement; 1 for the others.
int sweepBlockOf(const ifl_config &c, int dim, int slot) {
    if (c.smoother != IFL_SMOOTHER_RBGS) return 1;
    if (dim != 2) return (slot == 3 || slot == 4) ? std::min(std::max(c.sweep_block, 1), 2) : 1;
    if (slot < 3 || slot == ifl::kSlotT) {
        const int k = c.sweep_block_mom > 0 ? c.sweep_block_mom : std::min(std::max(c.sweep_block, 1), 2);
        return std::max(k, 1);
    }
    return std::max(c.sweep_block, 1);
}

// Whether an outer iteration's momentum call has an odd launch count for some component, so that
// it starts in the partner buffers (SmoothMode::fast), which must then equal the fields.
bool momentumStartsInPartner(const ifl_config &c, int dim) {
    for (int q = 0; q < dim; ++q) {
        const int n = c.n_iter[q], K = sweepBlockOf(c, dim, q);
        if (n > 0 && ((n + K - 1) / K) % 2 == 1) return true;
    }
    return false;
}

// Whether an outer iteration's p' smoother call (the reference's sweeps) has an odd launch count,
// so that it starts in the partner buffer, which k_pcorr_assemble then clears as well.
bool pressureStartsInPartner(const ifl_config &c, int dim) {
    if (c.p_solver != IFL_PSOLVE_SWEEPS) return false;
    const int n = c.n_iter[3], K = sweepBlockOf(c, dim, 3);
    return ((n + K - 1) / K) % 2 == 1;
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


// Smoother calls of Solver.cu:178-196.  Every launch reads one buffer of a ping-pong pair and
// writes the other.  Equations listed together share the matrix and are swept in the same
// launches (u and v in 2D); each keeps its own norms and stop state.
void enqueueSmoothers(const Plan &pl, const int *slots, int neq, const SmoothMode &md) {
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

void enqueueMomentumSmoothers(const Plan &pl, const SmoothMode &md) {
    if (pl.cfg->smoother == IFL_SMOOTHER_RBGS) {      // u, v (and w) share the matrix: one sweep kernel
        const int slots[3] = {0, 1, 2};
        enqueueSmoothers(pl, slots, pl.dim, md);
        return;
    }
    for (int q = 0; q < pl.dim; ++q) enqueueSmoothers(pl, &q, 1, md);   // :178, :180, :183
}

// The temperature sweeps (energy on), after the momentum sweeps on the same face velocities.
void enqueueTemperatureSmoother(const Plan &pl, const SmoothMode &md) {
    if (!pl.cfg->energy) return;
    const int slot = ifl::kSlotT;
    enqueueSmoothers(pl, &slot, 1, md);
}

// Solver.cu:196 plus the FP64-defect refinement cycles.  Returns the calls still to be closed
// (fast path) by the next kernel.
ifl::Finish enqueuePressureCorrection(const Plan &pl, bool fast) {
    if (pl.cfg->p_solver == IFL_PSOLVE_MG && (pl.mg || pl.mg3)) {    // mgEnsure made the levels
        if (pl.dim == 3) ifl::launch_mg3_solve(pl.L, *pl.mg3);
        else ifl::launch_mg_solve(pl.L, *pl.mg);
        return ifl::Finish{};
    }
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
    ifl::launch_pcorr_assemble(pl.L, pressureStartsInPartner(*pl.cfg, pl.dim));   // :189-193
    const ifl::Finish fp = enqueuePressureCorrection(pl, true);      // :196
    ifl::launch_update(pl.L, fp, momentumStartsInPartner(*pl.cfg, pl.dim));   // :198
    ifl::launch_reduce_cells(pl.L, ifl::RED_MOM | ifl::RED_MASS | ifl::RED_UPD | ifl::RED_FINISH);  // :200-268
}

size_t arenaLayout(ifl_solver *h, char *base, const bool *external) {
    ifl::CaseParams &P = h->P;
    const size_t N = (size_t)P.N, nx = P.nx, ny = P.ny, nz = P.nz;
    const size_t ncoef = N * (1 + 2 * P.dim);
    const size_t pstride = (size_t)partStride(P), ncb = (size_t)ifl::cell_blocks(P);
    P.partStride = (int)pstride;
    struct Part { void **dst; size_t bytes; };
    const size_t nT = h->cfg.energy ? N : 0;     // temperature matrix, source and partner: energy only
    auto own = [&](int q) { return external && external[q] ? size_t(0) : N * 4; };
    const Part parts[] = {
        {(void **)&P.u, own(0)}, {(void **)&P.v, own(1)}, {(void **)&P.w, own(2)}, {(void **)&P.p, own(3)},
        {(void **)&P.temp, own(4)}, {(void **)&P.pc, N * 4},
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
    size_t off = 0;
    for (const Part &p : parts) {
        if (base) *p.dst = p.bytes ? base + off : nullptr;
        off += align256(p.bytes);
    }
    return off;
}

size_t mgLayout(const ifl_config &c, MgHost &m, char *base) {
    if (c.dim == 3) {
        ifl::MgLaunch3 &G = m.launch3;
        int dims[ifl::kMgMaxLevels][3];
        G.nl = ifl::mg3_levels(c.nx, c.ny, c.nz, dims);
        G.lb = ifl::mg3_bottom_level(G.nl, dims);
        size_t off = 0;
        for (int l = 0; l < G.nl; ++l) {
            ifl::MgLevel3 &L = G.lv[l];
            L.nx = dims[l][0];
            L.ny = dims[l][1];
            L.nz = dims[l][2];
            if (l == 0) continue;
            float **pl[8] = {&L.x, &L.xa, &L.b, &L.C, &L.E, &L.N, &L.T, &L.dg};
            const size_t b = align256((size_t)L.nx * L.ny * L.nz * 4);
            for (float **q : pl) {
                *q = base ? (float *)(base + off) : nullptr;
                off += b;
            }
        }
        return off ? off : 256;       // a grid too small to coarsen has level 0 only
    }
    ifl::MgLaunch &G = m.launch;
    int dims[ifl::kMgMaxLevels][2];
    G.nl = ifl::mg_levels(c.nx, c.ny, dims);
    G.lb = ifl::mg_bottom_level(G.nl, dims);
    size_t off = align256(sizeof(ifl::MgLevel) * ifl::kMgMaxLevels);
    for (int l = 0; l < G.nl; ++l) {
        ifl::MgLevel &L = G.lv[l];
        L.nx = dims[l][0];
        L.ny = dims[l][1];
        if (l == 0) continue;
        float **pl[7] = {&L.x, &L.xu, &L.b, &L.C, &L.E, &L.N, &L.dg};
        const size_t b = align256((size_t)L.nx * L.ny * 4);
        for (float **q : pl) {
            *q = base ? (float *)(base + off) : nullptr;
            off += b;
        }
    }
    G.dev = (const ifl::MgLevel *)base;
    return off;
}

void mgSettings(const ifl_config &c, MgHost &m) {
    auto set = [&](auto &G) {
        G.cycles = c.mg_cycles > 0 ? c.mg_cycles : 1;
        G.nu = c.mg_sweeps > 0 ? c.mg_sweeps : 2;
        G.ncoarse = c.mg_coarse_sweeps > 0 ? c.mg_coarse_sweeps : 4;
        G.over = (float)(c.mg_over > 0.0 ? c.mg_over : 1.0);
    };
    set(m.launch);
    set(m.launch3);
}

Plan planOf(ifl_solver *h) {
    Plan pl;
    pl.L.P = &h->P;
    pl.L.batch = false;
    pl.L.tile = h->tile ? h->tile : ifl::TILE_64x16;
    pl.L.s = h->stream;
    pl.cfg = &h->cfg;
    pl.dim = h->P.dim;
    const bool mg = h->cfg.p_solver == IFL_PSOLVE_MG && h->mg;
    pl.mg = mg && pl.dim == 2 ? &h->mg->launch : nullptr;
    pl.mg3 = mg && pl.dim == 3 ? &h->mg->launch3 : nullptr;
    return pl;
}

int mgEnsure(ifl_solver *h) {
    const ifl_config &c = h->cfg;
    if (c.p_solver != IFL_PSOLVE_MG) return IFL_OK;
    if (!h->mg) {
        MgHost *m = new (std::nothrow) MgHost();
        if (!m) return fail(IFL_ERR_NOMEM, "out of host memory");
        const size_t total = mgLayout(c, *m, nullptr);
        hipError_t e = hipMalloc((void **)&m->arena, total);
        if (e != hipSuccess) {
            delete m;
            return fail(IFL_ERR_DEVICE, std::string("multigrid levels: ") + hipGetErrorString(e));
        }
        mgLayout(c, *m, m->arena);
        h->mg = m;
        e = hipMemsetAsync(m->arena, 0, total, h->stream);
        if (e == hipSuccess && c.dim == 2)
            e = hipMemcpyAsync(m->arena, m->launch.lv, sizeof(ifl::MgLevel) * ifl::kMgMaxLevels,
                               hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return fail(IFL_ERR_DEVICE, std::string("multigrid levels: ") + hipGetErrorString(e));
    }
    mgSettings(c, *h->mg);
    return IFL_OK;
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
    const int sm = mgEnsure(h);
    if (sm != IFL_OK) return sm;
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
    // an outer iteration whose momentum call starts in the partner buffers needs them equal to
    // the fields; the iterations keep them so (k_update's writeAlt), a call starting now sets them
    if (momentumStartsInPartner(h->cfg, h->P.dim)) {
        Plan pl = planOf(h);
        pl.L.s = s;
        for (int q = 0; q < h->P.dim; ++q) ifl::launch_copy_back(pl.L, q, 1);
        IFL_HIP(hipGetLastError());
    }
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

int finishSolve(ifl_solver *h, long long it0, const SolveTally &t, const ifl::CaseState &s, double ms,
                ifl_report *report) {
    if (report) {
        report->iterations = (int)(h->iter - it0);
        report->converged = s.done == 1;
        report->status = s.done == 2 ? IFL_ERR_DIVERGED : IFL_OK;
        report->elapsed_ms = ms;
        report->final_scaled = s.lastScaled;
        report->sweeps_total = t.sweeps;
    }
    if (s.done == 2) return fail(IFL_ERR_DIVERGED, "non-finite residual: the solve diverged");
    if (h->cfg.fail_if_not_converged && s.done != 1)
        return fail(IFL_ERR_NOT_CONVERGED, "the stop rule was not met in " + std::to_string(h->iter - it0)
                                               + " outer iterations");
    return IFL_OK;
}

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

}  // namespace ifl_host

using namespace ifl_host;

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
    const size_t total = arenaLayout(h, nullptr, nullptr);
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
    arenaLayout(h, h->arena, nullptr);
    ifl::CaseParams &P = h->P;
    const size_t N = (size_t)P.N;
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
    if (h->mg) {
        if (h->mg->arena) (void)hipFree(h->mg->arena);
        delete h->mg;
    }
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
    const int chunk = h->cfg.check_every > 0 ? std::min(h->cfg.check_every, kHistCap) : 64;
    std::vector<ifl_iter_record> ring((size_t)chunk);
    IFL_HIP(hipEventRecord(h->ev0, h->stream));
    if (h->iter == 0) ifl::launch_init_faces(planOf(h).L);           // Solver.cu:162-164
    ifl::CaseState s{};
    auto replay = [h](int n) -> int {
        for (int q = 0; q < n; ++q) IFL_HIP(hipGraphLaunch(h->graph, h->stream));
        return IFL_OK;
    };
    st = runChunks(h, max_iter, replay, t, ring.data(), chunk, s);
    if (st != IFL_OK) return st;
    IFL_HIP(hipEventRecord(h->ev1, h->stream));
    IFL_HIP(hipEventSynchronize(h->ev1));
    float ms = 0;
    IFL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    return finishSolve(h, it0, t, s, ms, report);
}

int ifl_run_stage(ifl_solver *h, int stage, double *out, size_t n_out) {
    if (!h) return fail(IFL_ERR_ARG, "null handle");
    double tmp[8] = {0};
    int st = setDone(h, 0);
    if (st != IFL_OK) return st;
    if ((st = mgEnsure(h)) != IFL_OK) return st;
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
    if (kernel < 0 || kernel >= IFL_KERNEL_COUNT) return fail(IFL_ERR_ARG, "bad kernel id");
    if (kernel == IFL_KERNEL_TEMP_SWEEP && !h->P.energy) return fail(IFL_ERR_ARG, "the temperature sweep needs energy = 1");
    int st = setDone(h, 0);
    if (st != IFL_OK) return st;
    if ((st = mgEnsure(h)) != IFL_OK) return st;
    // The timed launches overwrite device state (fields, partner buffers, matrices, monitor): the
    // whole arena is restored afterwards, on every return path, so a later solve continues exactly
    // as if no kernel had been timed.
    Snapshot snap;
    if ((st = snapTake(h, snap)) != IFL_OK) return st;
    struct Restore {
        ifl_solver *h;
        const Snapshot &s;
        bool armed = true;
        ~Restore() {
            if (armed) (void)snapRestore(h, s);
        }
    } guard{h, snap};
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
            case IFL_KERNEL_UPDATE:      // with the partner-buffer writes of the outer iteration's launch
                ifl::launch_update(L, ifl::Finish{}, momentumStartsInPartner(h->cfg, pl.dim));
                break;
            case IFL_KERNEL_PSOLVE: enqueuePressureCorrection(pl); break;
            default: break;
        }
    };
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
    guard.armed = false;
    if ((st = snapRestore(h, snap)) != IFL_OK) return st;
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
