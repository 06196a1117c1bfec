// simple_oracle.hpp - CPU oracle for the steady SIMPLE loop (TEST INFRASTRUCTURE ONLY).
//
// This is the checker, not the product: only tests, the smoke check and the CPU-baseline
// leg of the benchmark load it.  It restates, stage by stage and with the same literal
// types and evaluation order, the arithmetic of IncompFlow-CUDA:
//   include/kernels.cuh:10-1005  (every kernel body)
//   src/Solver.cu:153-299        (stage order of one outer iteration, stop rule)
//   src/Solver.cu:369-966        (index range of every launch, per-stage relax/tol)
//   include/constants.h:7-36     (face/slot order, geometry)
// Real = float reproduces the reference's rounding (scalar = float, include/types.h:4; double
// literals promote sub-expressions exactly as in the CUDA source); Real = double is the FP64
// accuracy oracle.  Build with -ffp-contract=off so no multiply-add is fused.
//
// Positions taken on the reference's defects (each also pinned by a test):
//   D1 stop rule: IFL_STOP_FAITHFUL keeps Solver.cu:237-245 as written, IFL_STOP_FIXED fixes it.
//   D2/D5 norms: sums run over in-domain terms only, in FP64, in a fixed (row) order.
//   D3 corner race: BC rows are applied face after face in launch order E,W,N,S,T,B.
//   D4 chaotic GS: red-black (colour (i+j+k)&1, red=0 first), lexicographic or Jacobi.
//   D6 Jacobi pointer swap: the result always ends in the caller's field.
//   T1-T5 (3D): the intended index forms are used (c + N*slot, i-fastest wf, j in T4, top in T5).
//
// Portions restate IncompFlow-CUDA, Copyright (c) WanZijun271, MIT License (LICENSE:1-3):
// Permission is hereby granted, free of charge, to any person obtaining a copy of this
// software and associated documentation files, to deal in the Software without restriction,
// subject to the condition that the copyright notice and this permission notice be included
// in all copies or substantial portions of the Software.  THE SOFTWARE IS PROVIDED "AS IS",
// WITHOUT WARRANTY OF ANY KIND.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../include/incompflow.h"

namespace oracle {

enum { E = 0, W = 1, N = 2, S = 3, T = 4, B = 5 };                   // constants.h:7-12
enum { aC = 0, aE = 1, aW = 2, aN = 3, aS = 4, aT = 5, aB = 6 };     // constants.h:14-20
enum { wall = 0, inlet = 1, outlet = 2 };                            // constants.h:34-36

inline int normal_dir(int face) { return face / 2; }  // E/W -> x, N/S -> y, T/B -> z

// Planted mistakes for the tests that show a pin is sensitive (0 = none; test use only).  Each
// code breaks exactly one stage the way a plausible transcription slip would.
enum {
    FAULT_NONE = 0,
    FAULT_DEFECT_SIGN = 1,       // defect: the per-cell output is A phi - b instead of b - A phi
    FAULT_DEFECT_NO_WEST = 2,    // defect: the west link term is dropped
    FAULT_REFINE_SUBTRACT = 3,   // refinement: p' -= delta instead of p' += delta
    FAULT_INITFACE_RIGHT = 4,    // initFaceVel: the face takes the right cell's value, not the mean
    FAULT_V_MATRIX_ROW = 5,      // v matrix: the wall rows skip the centre update (breaks I1)
    FAULT_MG_PROLONG_SIGN = 6,   // multigrid: the coarse correction is subtracted
    FAULT_MG_COARSE_EDGE = 7     // multigrid: a coarse x-coupling sums the first fine column, not the last
};
inline int g_fault = FAULT_NONE;

// The oracle keeps the reference's three momentum matrices (Solver.h:28-30, copied at
// Solver.cu:167-171); IFL_FIELD_MCOEF is the u matrix, these ids read the v and w copies.
enum { ORC_FIELD_VCOEF = 100, ORC_FIELD_WCOEF = 101 };

template <class Real>
class Simple {
public:
    // ---- case (config.h:6-81, constants.h:22-28) ----
    int dim = 2, nx = 0, ny = 0, nz = 1;
    int64_t Nc = 0;
    Real xmin, xmax, ymin, ymax, zmin, zmax;
    Real dx, dy, dz, areaX, areaY, areaZ;
    Real density, mu;
    int bct[6];
    Real bcv[6][3];
    std::vector<Real> prof[6];
    int nIter[4];
    Real relax[4];
    Real tol[4];
    Real tolMass;
    int smoother = IFL_SMOOTHER_RBGS;
    int stopRule = IFL_STOP_FAITHFUL;
    double tolRes = 1e-5;
    int pRefine = 0;
    int sweepBlock = 1;      // sweeps per block: a met stop test ends the call at the block's end
    int sweepBlockMom = 1;   // the same for the momentum and temperature equations
    int numOuterIter = 1;
    // p' solver (build addition): the reference's sweeps, or aggregation-multigrid V-cycles
    int pSolver = IFL_PSOLVE_SWEEPS;
    int mgCycles = 1, mgSweeps = 2, mgCoarse = 4;
    Real mgOver = 1;
    // temperature (config.h:19-23, 45-63; declared but never solved by the reference): the
    // momentum discretisation with mu -> k / c_p and the TempBCs rows (build addition)
    bool energy = false;
    Real gam = 0;
    int tbct[6];
    Real tbcv[6];
    int nIterT = 0;
    Real relaxT = 0, tolT = 0;

    // ---- state (Solver.h:10-45) ----
    std::vector<Real> u, v, w, p, temp, pc, uf, vf, wf;
    std::vector<Real> ucoef, vcoef, wcoef, pcoef;
    std::vector<Real> bu, bv, bw, bp;
    std::vector<Real> tcoef, bt;      // temperature matrix and source (energy on)

    // ---- monitor ----
    int64_t iter = 0;  // outer iterations executed on this object
    double pNormMax = 0, uNormMax = 0, vNormMax = 0, wNormMax = 0;
    double ufNormMax = 0, vfNormMax = 0, wfNormMax = 0;
    double resMax[5] = {0, 0, 0, 0, 0};  // continuity, u, v, w, T over the first 5 iterations
    int lastSweeps = 0;
    double lastSweepSum = 0;

    std::string error;

    bool init(const ifl_config &c) {
        if (c.dim != 2 && c.dim != 3) return fail("dim must be 2 or 3");
        if (c.nx < 4 || c.ny < 4 || (c.dim == 3 && c.nz < 4) || (c.dim == 2 && c.nz != 1))
            return fail("grid too small (need nx,ny(,nz) >= 4; nz == 1 in 2D)");
        dim = c.dim; nx = c.nx; ny = c.ny; nz = c.nz;
        Nc = (int64_t)nx * ny * nz;
        xmin = (Real)c.xmin; xmax = (Real)c.xmax;
        ymin = (Real)c.ymin; ymax = (Real)c.ymax;
        zmin = (Real)c.zmin; zmax = (Real)c.zmax;
        // constants.h:22-28
        dx = (xmax - xmin) / (Real)nx;
        dy = (ymax - ymin) / (Real)ny;
        dz = (zmax - zmin) / (Real)nz;
        areaX = dy * dz;
        areaY = dz * dx;
        areaZ = dx * dy;
        density = (Real)c.density;
        mu = (Real)c.viscosity;
        for (int f = 0; f < 6; ++f) {
            bct[f] = c.bc_type[f];
            if (bct[f] < 0 || bct[f] > 2) return fail("bad bc_type");
            for (int q = 0; q < 3; ++q) bcv[f][q] = (Real)c.bc_val[f][q];
        }
        for (int q = 0; q < 4; ++q) {
            nIter[q] = c.n_iter[q];
            relax[q] = (Real)c.relax[q];
            tol[q] = (Real)c.tol[q];
        }
        tolMass = (Real)c.tol_mass;
        smoother = c.smoother;
        stopRule = c.stop_rule;
        tolRes = c.tol_residual;
        pRefine = c.p_refine;
        // sweeps per launch on the device (red-black): in 2D sweep_block for p', sweep_block_mom for
        // the momentum and temperature equations (0: the smaller of sweep_block and 2); in 3D
        // sweep_block, at most 2, for p' and 1 for the others
        const bool blocked = c.dim == 2 && c.smoother == IFL_SMOOTHER_RBGS;
        const bool rb3 = c.dim == 3 && c.smoother == IFL_SMOOTHER_RBGS;
        sweepBlock = (blocked && c.sweep_block > 1) ? c.sweep_block : (rb3 && c.sweep_block > 1) ? 2 : 1;
        sweepBlockMom = !blocked ? 1 : c.sweep_block_mom > 0 ? c.sweep_block_mom : std::min(sweepBlock, 2);
        numOuterIter = c.num_outer_iter;
        pSolver = c.p_solver;
        if (pSolver != IFL_PSOLVE_SWEEPS && pSolver != IFL_PSOLVE_MG) return fail("bad p_solver");
        mgCycles = c.mg_cycles > 0 ? c.mg_cycles : 1;
        mgSweeps = c.mg_sweeps > 0 ? c.mg_sweeps : 2;
        mgCoarse = c.mg_coarse_sweeps > 0 ? c.mg_coarse_sweeps : 4;
        mgOver = (Real)(c.mg_over > 0 ? c.mg_over : 1.0);
        if (pSolver == IFL_PSOLVE_MG && c.p_refine > 0) return fail("p_refine needs p_solver = IFL_PSOLVE_SWEEPS");
        energy = c.energy != 0;
        gam = (Real)c.conductivity / (Real)c.specific_heat;
        for (int f = 0; f < 6; ++f) {
            tbct[f] = c.temp_bc_type[f];
            tbcv[f] = (Real)c.temp_bc_val[f];
            if (energy && tbct[f] != IFL_TEMP_DIRICHLET && tbct[f] != IFL_TEMP_NEUMANN) return fail("bad temp_bc_type");
        }
        nIterT = c.n_iter_t;
        relaxT = (Real)c.relax_t;
        tolT = (Real)c.tol_t;
        // Solver.cu:13-104: uniform initial u,v,w,p,temp; everything else zero.
        u.assign(Nc, (Real)c.init[0]);
        v.assign(Nc, (Real)c.init[1]);
        w.assign(Nc, (Real)c.init[2]);
        p.assign(Nc, (Real)c.init[3]);
        temp.assign(Nc, (Real)c.init[4]);
        pc.assign(Nc, 0);
        uf.assign((int64_t)(nx + 1) * ny * nz, 0);
        vf.assign((int64_t)nx * (ny + 1) * nz, 0);
        wf.assign((int64_t)nx * ny * (nz + 1), 0);
        int64_t ncoef = Nc * (1 + 2 * dim);
        ucoef.assign(ncoef, 0); vcoef.assign(ncoef, 0); wcoef.assign(ncoef, 0); pcoef.assign(ncoef, 0);
        bu.assign(Nc, 0); bv.assign(Nc, 0); bw.assign(Nc, 0); bp.assign(Nc, 0);
        if (energy) { tcoef.assign(ncoef, 0); bt.assign(Nc, 0); }
        return true;
    }

    bool fail(const char *m) { error = m; return false; }

    // ---- indexing (kernels.cuh:30-38, 105-114) ----
    inline int64_t cid(int i, int j, int k) const { return i + (int64_t)nx * (j + (int64_t)ny * k); }
    inline int64_t ufi(int i, int j, int k) const { return i + (int64_t)(nx + 1) * (j + (int64_t)ny * k); }
    inline int64_t vfi(int i, int j, int k) const { return i + (int64_t)nx * (j + (int64_t)(ny + 1) * k); }
    inline int64_t wfi(int i, int j, int k) const { return i + (int64_t)nx * (j + (int64_t)ny * k); }
    inline int64_t slot(int64_t c, int a) const { return c + Nc * a; }

    // Boundary value of component comp on face f at in-face position q
    // (velBCs::val[f][comp], config.h:35-42, optionally replaced by a per-face profile).
    inline Real bcval(int f, int comp, int64_t q) const {
        if (comp == normal_dir(f) && !prof[f].empty()) return prof[f][q];
        return bcv[f][comp];
    }
    inline bool ghost_copy(int f) const { return bct[f] == wall || bct[f] == inlet; }

    // ===================== stages =====================

    // initFaceVelKernel, kernels.cuh:10-45; launch ranges Solver.cu:384-398.
    void initFaceVel() {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                for (int i = 1; i < nx; ++i)
                    uf[ufi(i, j, k)] = g_fault == FAULT_INITFACE_RIGHT ? u[cid(i, j, k)]
                                                                       : (u[cid(i - 1, j, k)] + u[cid(i, j, k)]) / 2;
                if (j >= 1)
                    for (int i = 0; i < nx; ++i) vf[vfi(i, j, k)] = (v[cid(i, j - 1, k)] + v[cid(i, j, k)]) / 2;
                if (dim == 3 && k >= 1)
                    for (int i = 0; i < nx; ++i) wf[wfi(i, j, k)] = (w[cid(i, j, k - 1)] + w[cid(i, j, k)]) / 2;
            }
    }

    // applyBCsToFaceVelKernel, kernels.cuh:47-95; Solver.cu:410-465.
    inline Real faceBC(int f, int64_t q, Real adjacent) const {
        if (bct[f] == wall) return 0;                         // kernels.cuh:87-88
        if (bct[f] == inlet) return bcval(f, normal_dir(f), q); // :89-90
        return adjacent;                                      // :91-92 outlet
    }
    void applyBCsToFaceVel() {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                int64_t q = j + (int64_t)ny * k;
                uf[ufi(0, j, k)] = faceBC(W, q, u[cid(0, j, k)]);
                uf[ufi(nx, j, k)] = faceBC(E, q, u[cid(nx - 1, j, k)]);
            }
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) {
                int64_t q = i + (int64_t)nx * k;
                vf[vfi(i, 0, k)] = faceBC(S, q, v[cid(i, 0, k)]);
                vf[vfi(i, ny, k)] = faceBC(N, q, v[cid(i, ny - 1, k)]);
            }
        if (dim == 3)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int64_t q = i + (int64_t)nx * j;
                    wf[wfi(i, j, 0)] = faceBC(B, q, w[cid(i, j, 0)]);
                    wf[wfi(i, j, nz)] = faceBC(T, q, w[cid(i, j, nz - 1)]);
                }
    }

    // calcMomLinkCoefKernel, kernels.cuh:97-134 (T2/T3 fixed: wf i-fastest, slot c+N*a),
    // then the device copies uCoef -> vCoef (-> wCoef), Solver.cu:167-171.
    void calcMomLinkCoef() {
        Real *coef = ucoef.data();
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int64_t c = cid(i, j, k);
                    Real ue = uf[ufi(i + 1, j, k)];
                    Real uw = uf[ufi(i, j, k)];
                    Real vn = vf[vfi(i, j + 1, k)];
                    Real vs = vf[vfi(i, j, k)];
                    coef[slot(c, aE)] = (density * std::min(ue, (Real)0) * areaX) - mu * areaX / dx;
                    coef[slot(c, aW)] = (density * std::min(-uw, (Real)0) * areaX) - mu * areaX / dx;
                    coef[slot(c, aN)] = (density * std::min(vn, (Real)0) * areaY) - mu * areaY / dy;
                    coef[slot(c, aS)] = (density * std::min(-vs, (Real)0) * areaY) - mu * areaY / dy;
                    coef[slot(c, aC)] = -(coef[slot(c, aE)] + coef[slot(c, aW)] + coef[slot(c, aN)] + coef[slot(c, aS)]);
                    if (dim == 3) {
                        Real wt = wf[wfi(i, j, k + 1)];
                        Real wb = wf[wfi(i, j, k)];
                        coef[slot(c, aT)] = (density * std::min(wt, (Real)0) * areaZ) - mu * areaZ / dz;
                        coef[slot(c, aB)] = (density * std::min(-wb, (Real)0) * areaZ) - mu * areaZ / dz;
                        coef[slot(c, aC)] += -(coef[slot(c, aT)] + coef[slot(c, aB)]);
                    }
                }
        vcoef = ucoef;
        if (dim == 3) wcoef = ucoef;
    }

    // calcInteriorMomSrcTermKernel kernels.cuh:136-178 and calcBCMomSrcTermKernel :180-250.
    inline Real momSrc(int d, int i, int j, int k) const {
        int n = d == 0 ? nx : d == 1 ? ny : nz;
        int idx = d == 0 ? i : d == 1 ? j : k;
        int lo = 2 * d + 1, hi = 2 * d;  // W/S/B and E/N/T
        Real area = d == 0 ? areaX : d == 1 ? areaY : areaZ;
        int64_t st = d == 0 ? 1 : d == 1 ? (int64_t)nx : (int64_t)nx * ny;
        int64_t c = cid(i, j, k);
        Real p_L, p_R;
        if (idx == 0) {                                   // :232-238
            p_L = ghost_copy(lo) ? p[c] : (Real)0;
            p_R = p[c + st];
        } else if (idx == n - 1) {                        // :239-246
            p_R = ghost_copy(hi) ? p[c] : (Real)0;
            p_L = p[c - st];
        } else {                                          // :163-174
            p_L = p[c - st];
            p_R = p[c + st];
        }
        return 0.5 * (p_L - p_R) * area;                  // :176, :248
    }
    void calcMomSrcTerm() {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int64_t c = cid(i, j, k);
                    bu[c] = momSrc(0, i, j, k);
                    bv[c] = momSrc(1, i, j, k);
                    if (dim == 3) bw[c] = momSrc(2, i, j, k);
                }
    }

    // applyBCsToMomEqKernel kernels.cuh:252-334, faces applied in launch order
    // E, W, N, S (T, B) (Solver.cu:583-608; D3 resolved sequentially; T4 fixed).
    void bcRow(int f, int64_t c, int64_t q) {
        static const int exSlot[6] = {aE, aW, aN, aS, aT, aB};
        static const int inSlot[6] = {aW, aE, aS, aN, aB, aT};
        int64_t iC = slot(c, aC), iIn = slot(c, inSlot[f]), iEx = slot(c, exSlot[f]);
        Real *coefs[3] = {ucoef.data(), vcoef.data(), wcoef.data()};
        Real *srcs[3] = {bu.data(), bv.data(), bw.data()};
        for (int comp = 0; comp < dim; ++comp) {
            Real *coef = coefs[comp];
            if (bct[f] == wall || bct[f] == inlet) {     // :304-320
                Real phib = bcval(f, comp, q);
                if (!(g_fault == FAULT_V_MATRIX_ROW && comp == 1)) coef[iC] -= 2 * coef[iEx];
                coef[iIn] += coef[iEx] / 3;
                srcs[comp][c] -= (Real)8 / 3 * coef[iEx] * phib;
                coef[iEx] = 0;
            } else {                                      // :321-331 outlet
                coef[iC] += coef[iEx];
                coef[iEx] = 0;
            }
        }
    }
    void applyBCsToMomEq() {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) bcRow(E, cid(nx - 1, j, k), j + (int64_t)ny * k);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) bcRow(W, cid(0, j, k), j + (int64_t)ny * k);
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) bcRow(N, cid(i, ny - 1, k), i + (int64_t)nx * k);
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) bcRow(S, cid(i, 0, k), i + (int64_t)nx * k);
        if (dim == 3) {
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) bcRow(T, cid(i, j, nz - 1), i + (int64_t)nx * j);
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) bcRow(B, cid(i, j, 0), i + (int64_t)nx * j);
        }
    }

    // Temperature matrix and source (build addition): calcMomLinkCoefKernel (kernels.cuh:116-131)
    // with mu -> k / c_p, no interior source, then per face in launch order E, W, N, S (T, B)
    // the row of applyBCsToMomEqKernel for a given value (kernels.cuh:304-320) on Dirichlet faces
    // and the zero-gradient row (:321-331) on Neumann faces.
    void tempBcRow(int f, int64_t c) {
        static const int exSlot[6] = {aE, aW, aN, aS, aT, aB};
        static const int inSlot[6] = {aW, aE, aS, aN, aB, aT};
        int64_t iC = slot(c, aC), iIn = slot(c, inSlot[f]), iEx = slot(c, exSlot[f]);
        if (tbct[f] == IFL_TEMP_DIRICHLET) {
            bt[c] -= (Real)8 / 3 * tcoef[iEx] * tbcv[f];
            tcoef[iC] -= 2 * tcoef[iEx];
            tcoef[iIn] += tcoef[iEx] / 3;
            tcoef[iEx] = 0;
        } else {
            tcoef[iC] += tcoef[iEx];
            tcoef[iEx] = 0;
        }
    }
    void assembleTemperature() {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int64_t c = cid(i, j, k);
                    Real ue = uf[ufi(i + 1, j, k)], uw = uf[ufi(i, j, k)];
                    Real vn = vf[vfi(i, j + 1, k)], vs = vf[vfi(i, j, k)];
                    tcoef[slot(c, aE)] = (density * std::min(ue, (Real)0) * areaX) - gam * areaX / dx;
                    tcoef[slot(c, aW)] = (density * std::min(-uw, (Real)0) * areaX) - gam * areaX / dx;
                    tcoef[slot(c, aN)] = (density * std::min(vn, (Real)0) * areaY) - gam * areaY / dy;
                    tcoef[slot(c, aS)] = (density * std::min(-vs, (Real)0) * areaY) - gam * areaY / dy;
                    tcoef[slot(c, aC)] = -(tcoef[slot(c, aE)] + tcoef[slot(c, aW)] + tcoef[slot(c, aN)] + tcoef[slot(c, aS)]);
                    if (dim == 3) {
                        Real wt = wf[wfi(i, j, k + 1)], wb = wf[wfi(i, j, k)];
                        tcoef[slot(c, aT)] = (density * std::min(wt, (Real)0) * areaZ) - gam * areaZ / dz;
                        tcoef[slot(c, aB)] = (density * std::min(-wb, (Real)0) * areaZ) - gam * areaZ / dz;
                        tcoef[slot(c, aC)] += -(tcoef[slot(c, aT)] + tcoef[slot(c, aB)]);
                    }
                    bt[c] = 0;
                    if (i == nx - 1) tempBcRow(E, c);
                    if (i == 0) tempBcRow(W, c);
                    if (j == ny - 1) tempBcRow(N, c);
                    if (j == 0) tempBcRow(S, c);
                    if (dim == 3) {
                        if (k == nz - 1) tempBcRow(T, c);
                        if (k == 0) tempBcRow(B, c);
                    }
                }
    }

    // Solver.cu:167-175: one momentum assembly (and the temperature equation's, energy on).
    void assembleMomentum() {
        calcMomLinkCoef();
        calcMomSrcTerm();
        applyBCsToMomEq();
        if (energy) assembleTemperature();
    }

    // ---- smoothers: GaussSeidelIterateKernel kernels.cuh:926-1005 / pointJacobiIterateKernel
    //      :845-924, one cell update; neighbours outside the grid are 0 (:939-974).
    inline Real cellUpdate(const Real *fld, const Real *coef, const Real *src, int i, int j, int k,
                           Real rlx, Real &dPhi) const {
        int64_t c = cid(i, j, k);
        Real phi_C = fld[c];
        Real phi_E = (i != nx - 1) ? fld[c + 1] : (Real)0;
        Real phi_W = (i != 0) ? fld[c - 1] : (Real)0;
        Real phi_N = (j != ny - 1) ? fld[c + nx] : (Real)0;
        Real phi_S = (j != 0) ? fld[c - nx] : (Real)0;
        Real newPhi = src[c] - coef[slot(c, aE)] * phi_E - coef[slot(c, aW)] * phi_W - coef[slot(c, aN)] * phi_N
                      - coef[slot(c, aS)] * phi_S;
        if (dim == 3) {
            int64_t pl = (int64_t)nx * ny;
            Real phi_T = (k != nz - 1) ? fld[c + pl] : (Real)0;
            Real phi_B = (k != 0) ? fld[c - pl] : (Real)0;
            newPhi -= coef[slot(c, aT)] * phi_T + coef[slot(c, aB)] * phi_B;
        }
        newPhi /= coef[slot(c, aC)];
        dPhi = rlx * (newPhi - phi_C);
        return phi_C + dPhi;
    }

    // One sweep; returns the FP64 sum of squared increments (row partials summed in row order).
    double sweep(Real *fld, const Real *coef, const Real *src, Real rlx, int mode) {
        const int rows = ny * nz;
        std::vector<double> part(rows, 0.0);
        if (mode == IFL_SMOOTHER_LEX) {
            for (int r = 0; r < rows; ++r) {
                int j = r % ny, k = r / ny;
                double s = 0;
                for (int i = 0; i < nx; ++i) {
                    Real d;
                    fld[cid(i, j, k)] = cellUpdate(fld, coef, src, i, j, k, rlx, d);
                    s += (double)d * (double)d;
                }
                part[r] = s;
            }
        } else if (mode == IFL_SMOOTHER_JACOBI) {
            std::vector<Real> old(fld, fld + Nc);
#pragma omp parallel for schedule(static)
            for (int r = 0; r < rows; ++r) {
                int j = r % ny, k = r / ny;
                double s = 0;
                for (int i = 0; i < nx; ++i) {
                    Real d;
                    fld[cid(i, j, k)] = cellUpdate(old.data(), coef, src, i, j, k, rlx, d);
                    s += (double)d * (double)d;
                }
                part[r] = s;
            }
        } else {  // red-black: colour (i+j+k)&1, red (0) then black (1)
            for (int color = 0; color < 2; ++color) {
#pragma omp parallel for schedule(static)
                for (int r = 0; r < rows; ++r) {
                    int j = r % ny, k = r / ny;
                    double s = 0;
                    for (int i = (color + j + k) & 1; i < nx; i += 2) {
                        Real d;
                        fld[cid(i, j, k)] = cellUpdate(fld, coef, src, i, j, k, rlx, d);
                        s += (double)d * (double)d;
                    }
                    part[r] += s;
                }
            }
        }
        double tot = 0;
        for (int r = 0; r < rows; ++r) tot += part[r];
        return tot;
    }

    // GaussSeidelIterate / pointJacobiIterate host loops, Solver.cu:868-966 (norms in FP64).
    // With sweepBlock = 1 this is the reference loop; with sweepBlock = k the test still runs
    // after every sweep but a met test ends the call only after the current block of k sweeps.
    int iterate(Real *fld, const Real *coef, const Real *src, int n, Real rlx, Real tl) {
        return iterateBlocked(fld, coef, src, n, rlx, tl, sweepBlock);
    }
    int iterateMomentum(Real *fld, const Real *coef, const Real *src, int n, Real rlx, Real tl) {
        return iterateBlocked(fld, coef, src, n, rlx, tl, sweepBlockMom);
    }
    int iterateBlocked(Real *fld, const Real *coef, const Real *src, int n, Real rlx, Real tl, int block) {
        double maxNorm = -1e20;
        int done = 0;
        bool stopPending = false;
        lastSweepSum = 0;
        for (int it = 0; it < n; ++it) {
            double s = sweep(fld, coef, src, rlx, smoother);
            ++done;
            lastSweepSum = s;
            double norm = std::sqrt(s / (double)Nc);
            maxNorm = std::max(norm, maxNorm);
            double relNorm = norm / (maxNorm + 1e-20);
            if (relNorm < (double)tl) stopPending = true;
            if (stopPending && (it + 1) % block == 0) break;
        }
        lastSweeps = done;
        return done;
    }

    // FP64 L2 norm of b - A phi (build addition: momentum residual before the sweeps and the
    // p' defect).  Terms are accumulated in the order C, E, W, N, S, T, B.
    double defect(const Real *fld, const Real *coef, const Real *src, Real *out) const {
        const int rows = ny * nz;
        std::vector<double> part(rows, 0.0);
        const int64_t pl = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
        for (int r = 0; r < rows; ++r) {
            int j = r % ny, k = r / ny;
            double s = 0;
            for (int i = 0; i < nx; ++i) {
                int64_t c = cid(i, j, k);
                double rr = (double)src[c];
                rr -= (double)coef[slot(c, aC)] * (double)fld[c];
                rr -= (double)coef[slot(c, aE)] * (double)((i != nx - 1) ? fld[c + 1] : (Real)0);
                if (g_fault != FAULT_DEFECT_NO_WEST)
                    rr -= (double)coef[slot(c, aW)] * (double)((i != 0) ? fld[c - 1] : (Real)0);
                rr -= (double)coef[slot(c, aN)] * (double)((j != ny - 1) ? fld[c + nx] : (Real)0);
                rr -= (double)coef[slot(c, aS)] * (double)((j != 0) ? fld[c - nx] : (Real)0);
                if (dim == 3) {
                    rr -= (double)coef[slot(c, aT)] * (double)((k != nz - 1) ? fld[c + pl] : (Real)0);
                    rr -= (double)coef[slot(c, aB)] * (double)((k != 0) ? fld[c - pl] : (Real)0);
                }
                if (out) out[c] = (Real)(g_fault == FAULT_DEFECT_SIGN ? -rr : rr);
                s += rr * rr;
            }
            part[r] = s;
        }
        double tot = 0;
        for (int r = 0; r < rows; ++r) tot += part[r];
        return std::sqrt(tot);
    }

    // interiorRhieChowInterpolateKernel kernels.cuh:336-397 and bcRhieChowInterpolateKernel
    // :399-500 for the face between cells L=c-st and R=c (index along d: idx = R's index).
    inline Real rhieChow(int d, int64_t c, int idx, const Real *vel, const Real *coef) const {
        int n = d == 0 ? nx : d == 1 ? ny : nz;
        int lo = 2 * d + 1, hi = 2 * d;
        Real area = d == 0 ? areaX : d == 1 ? areaY : areaZ;
        int64_t st = d == 0 ? 1 : d == 1 ? (int64_t)nx : (int64_t)nx * ny;
        int64_t LId = c - st, RId = c;
        Real aC_L = coef[slot(LId, aC)];
        Real aC_R = coef[slot(RId, aC)];
        Real u_L = vel[LId];
        Real u_R = vel[RId];
        Real p_L = p[LId], p_R = p[RId], p_LL, p_RR;
        if (idx == 1) {                                   // west/south/bottom face 1, :481-487
            p_LL = ghost_copy(lo) ? p_L : (Real)0;
            p_RR = p[RId + st];
        } else if (idx == n - 1) {                        // east/north/top face n-1, :488-495
            p_RR = ghost_copy(hi) ? p_R : (Real)0;
            p_LL = p[LId - st];
        } else {                                          // :363-389
            p_LL = p[LId - st];
            p_RR = p[RId + st];
        }
        return 0.5 * (u_L + u_R) + (p_R - p_LL) * area / (4 * aC_L) + (p_RR - p_L) * area / (4 * aC_R)
               + 0.5 * (1 / aC_L + 1 / aC_R) * (p_L - p_R) * area;   // :394-395, :497-498
    }
    void RhieChowInterpolate() {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                for (int i = 1; i <= nx - 1; ++i) uf[ufi(i, j, k)] = rhieChow(0, cid(i, j, k), i, u.data(), ucoef.data());
                if (j >= 1)
                    for (int i = 0; i < nx; ++i) vf[vfi(i, j, k)] = rhieChow(1, cid(i, j, k), j, v.data(), vcoef.data());
                if (dim == 3 && k >= 1)
                    for (int i = 0; i < nx; ++i) wf[wfi(i, j, k)] = rhieChow(2, cid(i, j, k), k, w.data(), wcoef.data());
            }
        applyBCsToFaceVel();   // Solver.cu:187
    }

    // calcPresCorrLinkCoefKernel kernels.cuh:502-540 (T3 fixed).  Boundary neighbour slots
    // are never written and keep their zero from construction (Solver.cu:81).
    void calcPresCorrLinkCoef() {
        const int64_t pl = (int64_t)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int64_t c = cid(i, j, k);
                    Real *pcf = pcoef.data();
                    const Real *uC = ucoef.data(), *vC = vcoef.data(), *wC = wcoef.data();
                    if (i < nx - 1) pcf[slot(c, aE)] = -0.5 * density * areaX * areaX * (1 / uC[c] + 1 / uC[c + 1]);
                    if (i > 0) pcf[slot(c, aW)] = -0.5 * density * areaX * areaX * (1 / uC[c] + 1 / uC[c - 1]);
                    if (j < ny - 1) pcf[slot(c, aN)] = -0.5 * density * areaY * areaY * (1 / vC[c] + 1 / vC[c + nx]);
                    if (j > 0) pcf[slot(c, aS)] = -0.5 * density * areaY * areaY * (1 / vC[c] + 1 / vC[c - nx]);
                    pcf[slot(c, aC)] = -(pcf[slot(c, aE)] + pcf[slot(c, aW)] + pcf[slot(c, aN)] + pcf[slot(c, aS)]);
                    if (dim == 3) {
                        if (k < nz - 1) pcf[slot(c, aT)] = -0.5 * density * areaZ * areaZ * (1 / wC[c] + 1 / wC[c + pl]);
                        if (k > 0) pcf[slot(c, aB)] = -0.5 * density * areaZ * areaZ * (1 / wC[c] + 1 / wC[c - pl]);
                        pcf[slot(c, aC)] += -(pcf[slot(c, aT)] + pcf[slot(c, aB)]);
                    }
                }
    }

    // calcPresCorrSrcTermKernel kernels.cuh:542-581; returns sum of b'^2 (FP64, in-domain only).
    double calcPresCorrSrcTerm() {
        const int rows = ny * nz;
        std::vector<double> part(rows, 0.0);
#pragma omp parallel for schedule(static)
        for (int r = 0; r < rows; ++r) {
            int j = r % ny, k = r / ny;
            double s = 0;
            for (int i = 0; i < nx; ++i) {
                int64_t c = cid(i, j, k);
                Real b = density * (areaX * (uf[ufi(i, j, k)] - uf[ufi(i + 1, j, k)])
                                    + areaY * (vf[vfi(i, j, k)] - vf[vfi(i, j + 1, k)]));
                if (dim == 3) b += density * areaZ * (wf[wfi(i, j, k)] - wf[wfi(i, j, k + 1)]);
                bp[c] = b;
                s += (double)b * (double)b;
            }
            part[r] = s;
        }
        double tot = 0;
        for (int r = 0; r < rows; ++r) tot += part[r];
        return tot;
    }

    // ---- aggregation multigrid for the p' system (build addition, IFL_PSOLVE_MG) ----
    // Level 0 is the assembled p' matrix (kernels.cuh:502-540) plus, in cells next to an outlet
    // face, the coupling rho A^2 / aC to the ghost value p' = 0 that the velocity correction uses
    // there (kernels.cuh:713-727); level l+1 joins the 2x2(x2) blocks of level l (the last block of
    // an odd side holds one cell), while every side of level l is at least 4.  A coarse coupling is
    // the sum of the fine couplings across the coarse face, the coarse outlet term the sum of its
    // children's, and the coarse centre -(E + W + N + S) (-(T + B)) + outlet term, so the coarse
    // matrices keep the fine ones' symmetry and zero row sums.  Smoothing is unrelaxed red-black
    // Gauss-Seidel, x = (b - sum a_nb x_nb) / aC; the residual is FP64, restricted by summing the
    // children (FP64, order k, j, i) and rounded once; the coarse correction is x += over * x_c.
    struct MgLevel {
        int nx = 0, ny = 0, nz = 0;
        int64_t n = 0;
        std::vector<Real> x, b, C, E, N, T, dg;
    };
    std::vector<MgLevel> mg;

    static bool mgCoarsens(int nx_, int ny_, int nz_, int dim_) {
        return nx_ >= 4 && ny_ >= 4 && (dim_ == 2 || nz_ >= 4);
    }
    // Outlet term of cell c (faces in order E, W, N, S, T, B).
    Real mgOutletTerm(int i, int j, int k, int64_t c) const {
        Real dg = 0;
        const Real rinv = (Real)1 / ucoef[slot(c, aC)];
        const bool at[6] = {i == nx - 1, i == 0, j == ny - 1, j == 0, k == nz - 1, k == 0};
        for (int f = 0; f < 2 * dim; ++f) {
            if (!at[f] || bct[f] != outlet) continue;
            const Real A = f < 2 ? areaX : f < 4 ? areaY : areaZ;
            dg += density * A * A * rinv;
        }
        return dg;
    }
    void mgBuild() {
        mg.clear();
        MgLevel l0;
        l0.nx = nx; l0.ny = ny; l0.nz = nz; l0.n = Nc;
        l0.x.assign(Nc, 0);
        l0.b.assign(bp.begin(), bp.end());
        l0.dg.assign(Nc, 0);
        l0.C.assign(Nc, 0);
        l0.E.assign(pcoef.begin() + Nc * aE, pcoef.begin() + Nc * (aE + 1));
        l0.N.assign(pcoef.begin() + Nc * aN, pcoef.begin() + Nc * (aN + 1));
        if (dim == 3) l0.T.assign(pcoef.begin() + Nc * aT, pcoef.begin() + Nc * (aT + 1));
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int64_t c = cid(i, j, k);
                    l0.dg[c] = mgOutletTerm(i, j, k, c);
                    l0.C[c] = pcoef[slot(c, aC)] + l0.dg[c];
                }
        mg.push_back(std::move(l0));
        while ((int)mg.size() < 16 && mgCoarsens(mg.back().nx, mg.back().ny, mg.back().nz, dim)) {
            const MgLevel &F = mg.back();
            MgLevel Cl;
            Cl.nx = (F.nx + 1) / 2; Cl.ny = (F.ny + 1) / 2; Cl.nz = dim == 3 ? (F.nz + 1) / 2 : 1;
            Cl.n = (int64_t)Cl.nx * Cl.ny * Cl.nz;
            Cl.x.assign(Cl.n, 0); Cl.b.assign(Cl.n, 0); Cl.C.assign(Cl.n, 0);
            Cl.E.assign(Cl.n, 0); Cl.N.assign(Cl.n, 0); Cl.dg.assign(Cl.n, 0);
            if (dim == 3) Cl.T.assign(Cl.n, 0);
            auto fid = [&](int i, int j, int k) { return i + (int64_t)F.nx * (j + (int64_t)F.ny * k); };
            for (int K = 0; K < Cl.nz; ++K)
                for (int J = 0; J < Cl.ny; ++J)
                    for (int I = 0; I < Cl.nx; ++I) {
                        const int64_t C = I + (int64_t)Cl.nx * (J + (int64_t)Cl.ny * K);
                        const int i0 = 2 * I, i1 = std::min(2 * I + 1, F.nx - 1);
                        const int j0 = 2 * J, j1 = std::min(2 * J + 1, F.ny - 1);
                        const int k0 = dim == 3 ? 2 * K : 0, k1 = dim == 3 ? std::min(2 * K + 1, F.nz - 1) : 0;
                        Real e = 0, n = 0, t = 0, d = 0;
                        const int ie = g_fault == FAULT_MG_COARSE_EDGE ? i0 : i1;
                        for (int k = k0; k <= k1; ++k)
                            for (int j = j0; j <= j1; ++j) e += F.E[fid(ie, j, k)];
                        for (int k = k0; k <= k1; ++k)
                            for (int i = i0; i <= i1; ++i) n += F.N[fid(i, j1, k)];
                        if (dim == 3)
                            for (int j = j0; j <= j1; ++j)
                                for (int i = i0; i <= i1; ++i) t += F.T[fid(i, j, k1)];
                        for (int k = k0; k <= k1; ++k)
                            for (int j = j0; j <= j1; ++j)
                                for (int i = i0; i <= i1; ++i) d += F.dg[fid(i, j, k)];
                        Cl.E[C] = e; Cl.N[C] = n; Cl.dg[C] = d;
                        if (dim == 3) Cl.T[C] = t;
                    }
            const int64_t pl = (int64_t)Cl.nx * Cl.ny;
            for (int K = 0; K < Cl.nz; ++K)
                for (int J = 0; J < Cl.ny; ++J)
                    for (int I = 0; I < Cl.nx; ++I) {
                        const int64_t C = I + (int64_t)Cl.nx * (J + (int64_t)Cl.ny * K);
                        const Real w = I > 0 ? Cl.E[C - 1] : (Real)0, s_ = J > 0 ? Cl.N[C - Cl.nx] : (Real)0;
                        Real cc = -(Cl.E[C] + w + Cl.N[C] + s_);
                        if (dim == 3) {
                            const Real b_ = K > 0 ? Cl.T[C - pl] : (Real)0;
                            cc += -(Cl.T[C] + b_);
                        }
                        Cl.C[C] = cc + Cl.dg[C];
                    }
            mg.push_back(std::move(Cl));
        }
    }
    // n unrelaxed red-black sweeps of level L (colour (i+j+k)&1, red first).
    void mgSmooth(MgLevel &L, int n) {
        const int64_t pl = (int64_t)L.nx * L.ny;
        const int rows = L.ny * L.nz;
        for (int s = 0; s < n; ++s)
            for (int col = 0; col < 2; ++col) {
#pragma omp parallel for schedule(static)
                for (int r = 0; r < rows; ++r) {
                    const int j = r % L.ny, k = r / L.ny;
                    for (int i = (col + j + k) & 1; i < L.nx; i += 2) {
                        const int64_t c = i + (int64_t)L.nx * r;
                        const Real xE = i < L.nx - 1 ? L.x[c + 1] : (Real)0, xW = i > 0 ? L.x[c - 1] : (Real)0;
                        const Real xN = j < L.ny - 1 ? L.x[c + L.nx] : (Real)0, xS = j > 0 ? L.x[c - L.nx] : (Real)0;
                        const Real aW_ = i > 0 ? L.E[c - 1] : (Real)0, aS_ = j > 0 ? L.N[c - L.nx] : (Real)0;
                        Real v = L.b[c] - L.E[c] * xE - aW_ * xW - L.N[c] * xN - aS_ * xS;
                        if (dim == 3) {
                            const Real xT = k < L.nz - 1 ? L.x[c + pl] : (Real)0, xB = k > 0 ? L.x[c - pl] : (Real)0;
                            const Real aB_ = k > 0 ? L.T[c - pl] : (Real)0;
                            v -= L.T[c] * xT + aB_ * xB;
                        }
                        L.x[c] = v / L.C[c];
                    }
                }
            }
    }
    // FP64 residual of level L at cell (i, j, k), terms in the order C, E, W, N, S, T, B.
    double mgResidual(const MgLevel &L, int i, int j, int k) const {
        const int64_t pl = (int64_t)L.nx * L.ny, c = i + (int64_t)L.nx * (j + (int64_t)L.ny * k);
        double r = (double)L.b[c];
        r -= (double)L.C[c] * (double)L.x[c];
        r -= (double)L.E[c] * (double)(i < L.nx - 1 ? L.x[c + 1] : (Real)0);
        r -= (double)(i > 0 ? L.E[c - 1] : (Real)0) * (double)(i > 0 ? L.x[c - 1] : (Real)0);
        r -= (double)L.N[c] * (double)(j < L.ny - 1 ? L.x[c + L.nx] : (Real)0);
        r -= (double)(j > 0 ? L.N[c - L.nx] : (Real)0) * (double)(j > 0 ? L.x[c - L.nx] : (Real)0);
        if (dim == 3) {
            r -= (double)L.T[c] * (double)(k < L.nz - 1 ? L.x[c + pl] : (Real)0);
            r -= (double)(k > 0 ? L.T[c - pl] : (Real)0) * (double)(k > 0 ? L.x[c - pl] : (Real)0);
        }
        return r;
    }
    void mgRestrict(const MgLevel &F, MgLevel &Cl) {
#pragma omp parallel for schedule(static)
        for (int R = 0; R < Cl.ny * Cl.nz; ++R) {
            const int J = R % Cl.ny, K = R / Cl.ny;
            for (int I = 0; I < Cl.nx; ++I) {
                const int i1 = std::min(2 * I + 1, F.nx - 1), j1 = std::min(2 * J + 1, F.ny - 1);
                const int k0 = dim == 3 ? 2 * K : 0, k1 = dim == 3 ? std::min(2 * K + 1, F.nz - 1) : 0;
                double acc = 0.0;
                for (int k = k0; k <= k1; ++k)
                    for (int j = 2 * J; j <= j1; ++j)
                        for (int i = 2 * I; i <= i1; ++i) acc += mgResidual(F, i, j, k);
                const int64_t C = I + (int64_t)Cl.nx * R;
                Cl.b[C] = (Real)acc;
                Cl.x[C] = 0;
            }
        }
    }
    void mgProlong(MgLevel &F, const MgLevel &Cl) {
#pragma omp parallel for schedule(static)
        for (int r = 0; r < F.ny * F.nz; ++r) {
            const int j = r % F.ny, k = r / F.ny;
            for (int i = 0; i < F.nx; ++i) {
                const int64_t c = i + (int64_t)F.nx * r;
                const int64_t P_ = i / 2 + (int64_t)Cl.nx * (j / 2 + (int64_t)Cl.ny * (dim == 3 ? k / 2 : 0));
                F.x[c] = g_fault == FAULT_MG_PROLONG_SIGN ? F.x[c] - mgOver * Cl.x[P_] : F.x[c] + mgOver * Cl.x[P_];
            }
        }
    }
    void mgCycle(size_t l) {
        if (l + 1 == mg.size()) {
            mgSmooth(mg[l], mgCoarse);
            return;
        }
        mgSmooth(mg[l], mgSweeps);
        mgRestrict(mg[l], mg[l + 1]);
        mgCycle(l + 1);
        mgProlong(mg[l], mg[l + 1]);
        mgSmooth(mg[l], mgSweeps);
    }
    int solvePressureCorrectionMG() {
        mgBuild();
        for (int q = 0; q < mgCycles; ++q) mgCycle(0);
        std::copy(mg[0].x.begin(), mg[0].x.end(), pc.begin());
        lastSweeps = mgCycles * 2 * mgSweeps;
        lastSweepSum = 0;
        return lastSweeps;
    }

    // Solver.cu:193-196 p' = 0 and its smoother, followed by pRefine FP64-defect refinement
    // cycles (build addition): d = b' - A'p' in FP64, A'delta = (Real)d with the same
    // smoother from delta = 0, p' += delta.  Returns sweeps executed.
    int solvePressureCorrection() {
        std::fill(pc.begin(), pc.end(), (Real)0);
        if (pSolver == IFL_PSOLVE_MG) return solvePressureCorrectionMG();
        int sw = iterate(pc.data(), pcoef.data(), bp.data(), nIter[3], relax[3], tol[3]);
        double lastSum = lastSweepSum;
        if (pRefine > 0) {
            std::vector<Real> rr(Nc), delta(Nc);
            for (int r = 0; r < pRefine; ++r) {
                defect(pc.data(), pcoef.data(), bp.data(), rr.data());
                std::fill(delta.begin(), delta.end(), (Real)0);
                sw += iterate(delta.data(), pcoef.data(), rr.data(), nIter[3], relax[3], tol[3]);
                lastSum = lastSweepSum;
                if (g_fault == FAULT_REFINE_SUBTRACT)
                    for (int64_t c = 0; c < Nc; ++c) pc[c] -= delta[c];
                else
                    for (int64_t c = 0; c < Nc; ++c) pc[c] += delta[c];
            }
        }
        lastSweeps = sw;
        lastSweepSum = lastSum;
        return sw;
    }

    // updatePresKernel :814-843, updateInteriorVelKernel :583-649, updateBCVelKernel :651-748
    // (T5 fixed), updateFaceVelKernel :750-812.  sums[0..6] = p,u,v,w,uf,vf,wf.
    inline Real velCorr(int d, int i, int j, int k, const Real *coef) const {
        int n = d == 0 ? nx : d == 1 ? ny : nz;
        int idx = d == 0 ? i : d == 1 ? j : k;
        int lo = 2 * d + 1, hi = 2 * d;
        Real area = d == 0 ? areaX : d == 1 ? areaY : areaZ;
        int64_t st = d == 0 ? 1 : d == 1 ? (int64_t)nx : (int64_t)nx * ny;
        int64_t c = cid(i, j, k);
        Real pc_L, pc_R;
        if (idx == 0) {                                   // :713-719
            pc_L = ghost_copy(lo) ? pc[c] : (Real)0;
            pc_R = pc[c + st];
        } else if (idx == n - 1) {                        // :720-727
            pc_R = ghost_copy(hi) ? pc[c] : (Real)0;
            pc_L = pc[c - st];
        } else {                                          // :612-627
            pc_L = pc[c - st];
            pc_R = pc[c + st];
        }
        Real aC_C = coef[c];
        return relax_of(d) * 0.5 * (pc_L - pc_R) * area / aC_C;   // :630, :729
    }
    inline Real relax_of(int d) const { return relax[d]; }
    inline Real faceCorr(int d, int64_t c, const Real *coef) const {
        Real area = d == 0 ? areaX : d == 1 ? areaY : areaZ;
        int64_t st = d == 0 ? 1 : d == 1 ? (int64_t)nx : (int64_t)nx * ny;
        int64_t LId = c - st, RId = c;
        Real aC_L = coef[LId];
        Real aC_R = coef[RId];
        return relax_of(d) * 0.5 * (1 / aC_L + 1 / aC_R) * (pc[LId] - pc[RId]) * area;  // :793
    }
    void updateField(double sums[7]) {
        const int rows = ny * nz;
        std::vector<double> part((size_t)rows * 7, 0.0);
#pragma omp parallel for schedule(static)
        for (int r = 0; r < rows; ++r) {
            int j = r % ny, k = r / ny;
            double *s = &part[(size_t)r * 7];
            for (int i = 0; i < nx; ++i) {
                int64_t c = cid(i, j, k);
                Real dp = relax[3] * pc[c];                // :825
                p[c] += dp;
                s[0] += (double)dp * dp;
                Real du = velCorr(0, i, j, k, ucoef.data());
                u[c] += du;
                s[1] += (double)du * du;
                Real dv = velCorr(1, i, j, k, vcoef.data());
                v[c] += dv;
                s[2] += (double)dv * dv;
                if (dim == 3) {
                    Real dw = velCorr(2, i, j, k, wcoef.data());
                    w[c] += dw;
                    s[3] += (double)dw * dw;
                }
                if (i >= 1) {
                    Real d = faceCorr(0, c, ucoef.data());
                    uf[ufi(i, j, k)] += d;
                    s[4] += (double)d * d;
                }
                if (j >= 1) {
                    Real d = faceCorr(1, c, vcoef.data());
                    vf[vfi(i, j, k)] += d;
                    s[5] += (double)d * d;
                }
                if (dim == 3 && k >= 1) {
                    Real d = faceCorr(2, c, wcoef.data());
                    wf[wfi(i, j, k)] += d;
                    s[6] += (double)d * d;
                }
            }
        }
        for (int q = 0; q < 7; ++q) sums[q] = 0;
        for (int r = 0; r < rows; ++r)
            for (int q = 0; q < 7; ++q) sums[q] += part[(size_t)r * 7 + q];
    }

    // Momentum (and temperature) residual norms before the sweeps (build addition).
    void momentumResidual(double res[4]) {
        res[0] = defect(u.data(), ucoef.data(), bu.data(), nullptr);
        res[1] = defect(v.data(), vcoef.data(), bv.data(), nullptr);
        res[2] = dim == 3 ? defect(w.data(), wcoef.data(), bw.data(), nullptr) : 0.0;
        res[3] = energy ? defect(temp.data(), tcoef.data(), bt.data(), nullptr) : 0.0;
    }

    // The temperature sweeps (build addition): the momentum smoother call on T.
    int smoothTemperature() { return iterateMomentum(temp.data(), tcoef.data(), bt.data(), nIterT, relaxT, tolT); }

    // ===================== one outer iteration, Solver.cu:160-290 =====================
    // Returns true when the stop rule holds.
    bool outerIteration(ifl_iter_record &rec) {
        std::memset(&rec, 0, sizeof(rec));
        if (iter == 0) initFaceVel();                    // :162-164
        applyBCsToFaceVel();                             // :165
        assembleMomentum();                              // :167-175
        double res[4];
        momentumResidual(res);
        rec.sweeps[0] = iterateMomentum(u.data(), ucoef.data(), bu.data(), nIter[0], relax[0], tol[0]);   // :178
        rec.sweeps[1] = iterateMomentum(v.data(), vcoef.data(), bv.data(), nIter[1], relax[1], tol[1]);   // :180
        if (dim == 3) rec.sweeps[2] = iterateMomentum(w.data(), wcoef.data(), bw.data(), nIter[2], relax[2], tol[2]);
        if (energy) rec.sweeps_t = smoothTemperature();
        RhieChowInterpolate();                           // :186-187
        calcPresCorrLinkCoef();                          // :189
        double mass = calcPresCorrSrcTerm();             // :191
        rec.sweeps[3] = solvePressureCorrection();       // :193-196
        double sums[7];
        updateField(sums);                               // :198
        return finishIteration(rec, sums, mass, res);
    }

    // Solver.cu:226-268 (norms in FP64) plus the scaled residual.
    bool finishIteration(ifl_iter_record &rec, const double sums[7], double mass, const double res[4]) {
        const double Nd = (double)Nc;
        rec.p_norm = std::sqrt(sums[0] / Nd);
        rec.u_norm = std::sqrt(sums[1] / Nd);
        rec.v_norm = std::sqrt(sums[2] / Nd);
        rec.uf_norm = std::sqrt(sums[4] / ((double)(nx - 1) * ny * nz));
        rec.vf_norm = std::sqrt(sums[5] / ((double)nx * (ny - 1) * nz));
        rec.mass_norm = std::sqrt(mass / Nd);
        if (dim == 3) {
            rec.w_norm = std::sqrt(sums[3] / Nd);
            rec.wf_norm = std::sqrt(sums[6] / ((double)nx * ny * (nz - 1)));
        }
        rec.res_cont = std::sqrt(mass);
        rec.res_u = res[0];
        rec.res_v = res[1];
        rec.res_w = res[2];
        rec.res_t = res[3];
        double rs[5] = {rec.res_cont, rec.res_u, rec.res_v, rec.res_w, rec.res_t};
        if (iter < 5)
            for (int q = 0; q < 5; ++q) resMax[q] = std::max(resMax[q], rs[q]);
        double sc = 0;
        for (int q = 0; q < 5; ++q)
            if ((q < 3 || (q == 3 && dim == 3) || (q == 4 && energy)) && resMax[q] > 0)
                sc = std::max(sc, rs[q] / resMax[q]);
        rec.scaled = sc;

        pNormMax = std::max(pNormMax, rec.p_norm);       // :237
        if (stopRule == IFL_STOP_FAITHFUL) {             // :238-245 as written (D1)
            uNormMax = vNormMax = ufNormMax = vfNormMax = wNormMax = wfNormMax = std::max(pNormMax, rec.p_norm);
        } else {
            uNormMax = std::max(uNormMax, rec.u_norm);
            vNormMax = std::max(vNormMax, rec.v_norm);
            wNormMax = std::max(wNormMax, rec.w_norm);
            ufNormMax = std::max(ufNormMax, rec.uf_norm);
            vfNormMax = std::max(vfNormMax, rec.vf_norm);
            wfNormMax = std::max(wfNormMax, rec.wf_norm);
        }
        rec.rel_p = rec.p_norm / pNormMax;               // :247-256
        rec.rel_u = rec.u_norm / uNormMax;
        rec.rel_v = rec.v_norm / vNormMax;
        rec.rel_uf = rec.uf_norm / ufNormMax;
        rec.rel_vf = rec.vf_norm / vfNormMax;
        if (dim == 3) {
            rec.rel_w = rec.w_norm / wNormMax;
            rec.rel_wf = rec.wf_norm / wfNormMax;
        }
        bool conv = false;
        if (stopRule == IFL_STOP_FAITHFUL || stopRule == IFL_STOP_FIXED) {   // :258-268
            conv = rec.rel_p < (double)tol[3] && rec.rel_u < (double)tol[0] && rec.rel_v < (double)tol[1]
                   && rec.rel_uf < (double)tol[0] && rec.rel_vf < (double)tol[1];
            if (dim == 3) conv = conv && rec.rel_w < (double)tol[2] && rec.rel_wf < (double)tol[2];
            conv = conv || rec.mass_norm < (double)tolMass;
        } else if (stopRule == IFL_STOP_SCALED) {
            conv = sc < tolRes;
        }
        rec.converged = conv ? 1 : 0;
        ++iter;
        return conv;
    }

    bool finiteRecord(const ifl_iter_record &r) const {
        const double v[] = {r.p_norm, r.u_norm, r.v_norm, r.w_norm, r.uf_norm, r.vf_norm, r.wf_norm,
                            r.mass_norm, r.res_cont, r.res_u, r.res_v, r.res_w, r.res_t};
        for (double x : v)
            if (!std::isfinite(x)) return false;
        return true;
    }

    // Solver::solve loop; returns iterations executed, sets *status.
    int solve(int maxIter, ifl_iter_record *hist, size_t cap, int *converged, int *status) {
        *converged = 0;
        *status = IFL_OK;
        int n = 0;
        for (int it = 0; it < maxIter; ++it) {
            ifl_iter_record rec;
            bool c = outerIteration(rec);
            if ((size_t)n < cap && hist) hist[n] = rec;
            ++n;
            if (!finiteRecord(rec)) { *status = IFL_ERR_DIVERGED; break; }
            if (c) { *converged = 1; break; }
        }
        return n;
    }

    // Field access by IFL_FIELD_* id.
    std::vector<Real> *field(int id) {
        switch (id) {
            case IFL_FIELD_U: return &u;
            case IFL_FIELD_V: return &v;
            case IFL_FIELD_W: return &w;
            case IFL_FIELD_P: return &p;
            case IFL_FIELD_TEMP: return &temp;
            case IFL_FIELD_PCORR: return &pc;
            case IFL_FIELD_UF: return &uf;
            case IFL_FIELD_VF: return &vf;
            case IFL_FIELD_WF: return &wf;
            case IFL_FIELD_MCOEF: return &ucoef;
            case IFL_FIELD_PCOEF: return &pcoef;
            case IFL_FIELD_BU: return &bu;
            case IFL_FIELD_BV: return &bv;
            case IFL_FIELD_BW: return &bw;
            case IFL_FIELD_BP: return &bp;
            case IFL_FIELD_TCOEF: return energy ? &tcoef : nullptr;
            case IFL_FIELD_BT: return energy ? &bt : nullptr;
            case ORC_FIELD_VCOEF: return &vcoef;
            case ORC_FIELD_WCOEF: return &wcoef;
            default: return nullptr;
        }
    }
};

}  // namespace oracle
