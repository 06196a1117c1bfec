// kernels.hip - gfx950 kernels of one SIMPLE outer iteration on a collocated structured grid.
//
// Cell kernels: a workgroup of 256 threads (4 wave64) covers 64 cells along x and 16 rows
// (4 rows per thread), so every wave reads and writes 256 contiguous bytes per row.  Each cell
// is done by one thread for all of its stage's work, so the boundary rows that the reference
// modifies from four concurrent streams (Solver.cu:583-597) are applied race-free and in launch
// order E,W,N,S,T,B.  FP64 norms are reduced with wave64 butterflies into one record per
// workgroup, and added up in a fixed order by k_reduce_cells (deterministic; replaces the float
// atomicAdd of kernels.cuh:578 etc.).  Arithmetic follows kernels.cuh expression by expression,
// with the same float/double promotions, and is compiled with -ffp-contract=off, so a stage's
// output equals the FP32 CPU oracle's bit for bit.
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_stencil.hpp"

namespace ifl {

// Visit the cells of this workgroup: f(i, j, k, c).
template <class F>
__device__ __forceinline__ void forCells(const CaseParams &P, const Blk &b, F &&f) {
    const int rows = P.ny * P.nz;
    const int i = b.x * 64 + (threadIdx.x & 63);
    const int r0 = b.y * (4 * kRows) + (threadIdx.x >> 6);
#pragma unroll
    for (int q = 0; q < kRows; ++q) {
        const int r = r0 + 4 * q;
        if (i < P.nx && r < rows) {
            const int k = r / P.ny;
            const int j = r - k * P.ny;
            f(i, j, k, i + P.nx * r);
        }
    }
}

// ---------------- initFaceVel, kernels.cuh:10-45 ----------------
template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_init_faces(CaseParams P0, BatchMap M) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    forCells(P, blk, [&](int i, int j, int k, int c) {
        if (i >= 1) P.uf[ufi(P, i, j, k)] = (P.u[c - 1] + P.u[c]) / 2;
        if (j >= 1) P.vf[vfi(P, i, j, k)] = (P.v[c - P.nx] + P.v[c]) / 2;
        if (DIM == 3 && k >= 1) P.wf[wfi(P, i, j, k)] = (P.w[c - P.nx * P.ny] + P.w[c]) / 2;
    });
}

// ---------------- applyBCsToFaceVel, kernels.cuh:47-95 ----------------
template <int DIM>
__device__ __forceinline__ void applyFaceBCs(const CaseParams &P, int i, int j, int k, int c) {
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    if (i == 0) P.uf[ufi(P, 0, j, k)] = faceBC(P, W, j + ny * k, P.u[c]);
    if (i == nx - 1) P.uf[ufi(P, nx, j, k)] = faceBC(P, E, j + ny * k, P.u[c]);
    if (j == 0) P.vf[vfi(P, i, 0, k)] = faceBC(P, S, i + nx * k, P.v[c]);
    if (j == ny - 1) P.vf[vfi(P, i, ny, k)] = faceBC(P, N, i + nx * k, P.v[c]);
    if (DIM == 3) {
        if (k == 0) P.wf[wfi(P, i, j, 0)] = faceBC(P, B, i + nx * j, P.w[c]);
        if (k == nz - 1) P.wf[wfi(P, i, j, nz)] = faceBC(P, T, i + nx * j, P.w[c]);
    }
}

template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_face_bc(CaseParams P0, BatchMap M) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    forCells(P, blk, [&](int i, int j, int k, int c) { applyFaceBCs<DIM>(P, i, j, k, c); });
}

// ---------------- momentum assembly ----------------
// calcMomLinkCoefKernel (kernels.cuh:97-134), calcInteriorMomSrcTermKernel (:136-178),
// calcBCMomSrcTermKernel (:180-250) and applyBCsToMomEqKernel (:252-334) for one cell, with the
// start-of-iteration face BCs (Solver.cu:165) applied first.  One coefficient set serves u, v, w
// (the reference's copies are bitwise equal, Solver.cu:167-171).  Also accumulates the FP64
// momentum residual sum (b - A phi)^2 of each component before the sweeps.
__device__ __forceinline__ float momSrc(const CaseParams &P, int d, int idx, int n, int c, int st) {
    const int lo = 2 * d + 1, hi = 2 * d;
    const float area = d == 0 ? P.Ax : d == 1 ? P.Ay : P.Az;
    float p_L, p_R;
    if (idx == 0) {
        p_L = ghostCopy(P, lo) ? P.p[c] : 0.0f;
        p_R = P.p[c + st];
    } else if (idx == n - 1) {
        p_R = ghostCopy(P, hi) ? P.p[c] : 0.0f;
        p_L = P.p[c - st];
    } else {
        p_L = P.p[c - st];
        p_R = P.p[c + st];
    }
    return 0.5 * (p_L - p_R) * area;
}

template <int DIM>
__device__ __forceinline__ void bcRow(const CaseParams &P, int f, int q, float (&a)[7], float (&b)[3]) {
    const float aEx = a[bcExSlot(f)];
    if (ghostCopy(P, f)) {                       // kernels.cuh:307, 312, 318
#pragma unroll
        for (int comp = 0; comp < DIM; ++comp) b[comp] -= (float)8 / 3 * aEx * bcval(P, f, comp, q);
    }
    bcRowCoef(P, f, a);
}

template <int DIM>
__device__ __forceinline__ double momResidual(const CaseParams &P, const float *fld, const float (&a)[7], float b,
                                              int i, int j, int k, int c) {
    const int pl = P.nx * P.ny;
    double r = (double)b;
    r -= (double)a[sC] * (double)fld[c];
    r -= (double)a[sE] * (double)((i != P.nx - 1) ? fld[c + 1] : 0.0f);
    r -= (double)a[sW] * (double)((i != 0) ? fld[c - 1] : 0.0f);
    r -= (double)a[sN] * (double)((j != P.ny - 1) ? fld[c + P.nx] : 0.0f);
    r -= (double)a[sS] * (double)((j != 0) ? fld[c - P.nx] : 0.0f);
    if (DIM == 3) {
        r -= (double)a[sT] * (double)((k != P.nz - 1) ? fld[c + pl] : 0.0f);
        r -= (double)a[sB] * (double)((k != 0) ? fld[c - pl] : 0.0f);
    }
    return r;
}

// Boundary row of face f of the temperature equation: the momentum rows with the face's
// temperature BC (Dirichlet value, or zero gradient).
template <int DIM>
__device__ __forceinline__ void tempBcRow(const CaseParams &P, int f, float (&a)[7], float &b) {
    const bool given = tempDirichlet(P, f);
    if (given) b -= (float)8 / 3 * a[bcExSlot(f)] * P.tbcv[f];
    bcRowCoefKind(given, f, a);
}

// The seven-point neighbourhood of a cell (C, E, W, N, S, T, B): offsets to the neighbours, 0
// (the cell itself) where the domain ends, so that every neighbour load is issued unconditionally
// and the missing ones are replaced by the caller.  With no branch around a load, the loads of a
// cell are all in flight together instead of one dependent round trip each.
struct Nbr {
    int off[7];
    bool has[7];
};
template <int DIM>
__device__ __forceinline__ Nbr nbrOf(const CaseParams &P, int i, int j, int k) {
    const int pl = P.nx * P.ny;
    Nbr n;
    n.has[0] = true;
    n.has[1] = i != P.nx - 1;
    n.has[2] = i != 0;
    n.has[3] = j != P.ny - 1;
    n.has[4] = j != 0;
    n.has[5] = DIM == 3 && k != P.nz - 1;
    n.has[6] = DIM == 3 && k != 0;
    const int d[7] = {0, 1, -1, P.nx, -P.nx, pl, -pl};
#pragma unroll
    for (int s = 0; s < 7; ++s) n.off[s] = n.has[s] ? d[s] : 0;
    return n;
}
template <int DIM>
__device__ __forceinline__ void gather7(const float *f, int c, const Nbr &n, float (&x)[7]) {
#pragma unroll
    for (int s = 0; s < 1 + 2 * DIM; ++s) x[s] = f[c + n.off[s]];
}
// momResidual on a gathered neighbourhood (the same terms in the same order).
template <int DIM>
__device__ __forceinline__ double residual7(const float (&a)[7], float b, const float (&x)[7], const Nbr &n) {
    double r = (double)b;
    r -= (double)a[sC] * (double)x[0];
    r -= (double)a[sE] * (double)(n.has[1] ? x[1] : 0.0f);
    r -= (double)a[sW] * (double)(n.has[2] ? x[2] : 0.0f);
    r -= (double)a[sN] * (double)(n.has[3] ? x[3] : 0.0f);
    r -= (double)a[sS] * (double)(n.has[4] ? x[4] : 0.0f);
    if (DIM == 3) {
        r -= (double)a[sT] * (double)(n.has[5] ? x[5] : 0.0f);
        r -= (double)a[sB] * (double)(n.has[6] ? x[6] : 0.0f);
    }
    return r;
}
// faceBC (kernels.cuh:87-93) with the face's profile value already loaded.
__device__ __forceinline__ float faceBCv(const CaseParams &P, int f, float prof, float adjacent) {
    const int t = P.bct[f];
    if (t == IFL_WALL) return 0.0f;
    if (t == IFL_INLET) return P.prof[f] != nullptr ? prof : P.bcv[f][f >> 1];
    return adjacent;
}
// momSrc on a gathered pressure neighbourhood: direction d, position idx of n cells, lo/hi
// neighbour slots.
__device__ __forceinline__ float momSrc7(const CaseParams &P, int d, int idx, int n, const float (&p)[7]) {
    const int lo = 2 * d + 1, hi = 2 * d;
    const float area = d == 0 ? P.Ax : d == 1 ? P.Ay : P.Az;
    const float p_L = (idx == 0) ? (ghostCopy(P, lo) ? p[0] : 0.0f) : p[2 + 2 * d];
    const float p_R = (idx == n - 1) ? (ghostCopy(P, hi) ? p[0] : 0.0f) : p[1 + 2 * d];
    return 0.5 * (p_L - p_R) * area;
}

template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_mom_assemble(CaseParams P0, BatchMap M) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    if (leader(blk))
        for (int q = 0; q < 5; ++q) P.st->sweepsIter[q] = 0;   // a new outer iteration starts
    double rs[4] = {0.0, 0.0, 0.0, 0.0};
    forCells(P, blk, [&](int i, int j, int k, int c) {
        const int nx = P.nx, ny = P.ny, nz = P.nz;
        const Nbr nb = nbrOf<DIM>(P, i, j, k);
        // ---- every load of the cell first ----
        float xu[7], xv[7], xw[7] = {0, 0, 0, 0, 0, 0, 0}, Pn[7];
        gather7<DIM>(P.u, c, nb, xu);
        gather7<DIM>(P.v, c, nb, xv);
        if (DIM == 3) gather7<DIM>(P.w, c, nb, xw);
        gather7<DIM>(P.p, c, nb, Pn);
        const float fE = P.uf[ufi(P, i + 1, j, k)], fW = P.uf[ufi(P, i, j, k)];
        const float fN = P.vf[vfi(P, i, j + 1, k)], fS = P.vf[vfi(P, i, j, k)];
        const float fT = DIM == 3 ? P.wf[wfi(P, i, j, k + 1)] : 0.0f, fB = DIM == 3 ? P.wf[wfi(P, i, j, k)] : 0.0f;
        // per-face profile values (read from u where a face has none; then unused)
        const int qf[6] = {j + ny * k, j + ny * k, i + nx * k, i + nx * k, i + nx * j, i + nx * j};
        float prv[6];
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f) prv[f] = (P.prof[f] != nullptr ? P.prof[f] : P.u)[qf[f]];
        // ---- face BCs of the boundary faces (Solver.cu:165), kernels.cuh:47-95 ----
        const float adj[6] = {xu[0], xu[0], xv[0], xv[0], xw[0], xw[0]};
        float fb[6];
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f) fb[f] = faceBCv(P, f, prv[f], adj[f]);
        if (i == 0) P.uf[ufi(P, 0, j, k)] = fb[W];
        if (i == nx - 1) P.uf[ufi(P, nx, j, k)] = fb[E];
        if (j == 0) P.vf[vfi(P, i, 0, k)] = fb[S];
        if (j == ny - 1) P.vf[vfi(P, i, ny, k)] = fb[N];
        if (DIM == 3) {
            if (k == 0) P.wf[wfi(P, i, j, 0)] = fb[B];
            if (k == nz - 1) P.wf[wfi(P, i, j, nz)] = fb[T];
        }
        // face velocities (boundary faces hold the values just applied)
        const float ue = (i == nx - 1) ? fb[E] : fE;
        const float uw = (i == 0) ? fb[W] : fW;
        const float vn = (j == ny - 1) ? fb[N] : fN;
        const float vs = (j == 0) ? fb[S] : fS;
        const float wt = DIM == 3 ? ((k == nz - 1) ? fb[T] : fT) : 0.0f;
        const float wb = DIM == 3 ? ((k == 0) ? fb[B] : fB) : 0.0f;
        float a[7];
        momLinkCoef<DIM>(P, ue, uw, vn, vs, wt, wb, a);
        float b[3];
        b[0] = momSrc7(P, 0, i, nx, Pn);
        b[1] = momSrc7(P, 1, j, ny, Pn);
        b[2] = DIM == 3 ? momSrc7(P, 2, k, nz, Pn) : 0.0f;
        // BC rows in launch order E, W, N, S, T, B (Solver.cu:583-608), kernels.cuh:304-331
        auto row = [&](int f) {
            const float aEx = a[bcExSlot(f)];
            if (ghostCopy(P, f)) {
#pragma unroll
                for (int comp = 0; comp < DIM; ++comp) {
                    const float val = (comp == (f >> 1) && P.prof[f] != nullptr) ? prv[f] : P.bcv[f][comp];
                    b[comp] -= (float)8 / 3 * aEx * val;
                }
            }
            bcRowCoef(P, f, a);
        };
        if (i == nx - 1) row(E);
        if (i == 0) row(W);
        if (j == ny - 1) row(N);
        if (j == 0) row(S);
        if (DIM == 3) {
            if (k == nz - 1) row(T);
            if (k == 0) row(B);
        }
#pragma unroll
        for (int s = 0; s < 1 + 2 * DIM; ++s) P.mcoef[c + P.N * s] = a[s];
        P.bu[c] = b[0];
        P.bv[c] = b[1];
        if (DIM == 3) P.bw[c] = b[2];
        const double r0 = residual7<DIM>(a, b[0], xu, nb);
        const double r1 = residual7<DIM>(a, b[1], xv, nb);
        rs[0] += r0 * r0;
        rs[1] += r1 * r1;
        if (DIM == 3) {
            const double r2 = residual7<DIM>(a, b[2], xw, nb);
            rs[2] += r2 * r2;
        }
        if (P.energy) {      // the temperature equation on the same face velocities
            float Tn[7];
            gather7<DIM>(P.temp, c, nb, Tn);
            float t[7];
            linkCoef<DIM>(P, P.gam, ue, uw, vn, vs, wt, wb, t);
            float bt = 0.0f;
            if (i == nx - 1) tempBcRow<DIM>(P, E, t, bt);
            if (i == 0) tempBcRow<DIM>(P, W, t, bt);
            if (j == ny - 1) tempBcRow<DIM>(P, N, t, bt);
            if (j == 0) tempBcRow<DIM>(P, S, t, bt);
            if (DIM == 3) {
                if (k == nz - 1) tempBcRow<DIM>(P, T, t, bt);
                if (k == 0) tempBcRow<DIM>(P, B, t, bt);
            }
#pragma unroll
            for (int s = 0; s < 1 + 2 * DIM; ++s) P.tcoef[c + P.N * s] = t[s];
            P.bt[c] = bt;
            const double r3 = residual7<DIM>(t, bt, Tn, nb);
            rs[3] += r3 * r3;
        }
    });
    storeBlockSums<4>(rs, P.partMom + 4 * blk.flat);
}

// ---------------- one relaxed Gauss-Seidel/Jacobi cell update ----------------
// GaussSeidelIterateKernel / pointJacobiIterateKernel (kernels.cuh:926-1005, 845-924) for one
// cell reading `fld`; neighbours outside the grid are 0 (:939-974).
template <int DIM>
__device__ __forceinline__ float gsUpdate(const CaseParams &P, const float *fld, const float *coef, const float *src,
                                          int i, int j, int k, int c, float rlx, float &dPhi) {
    const int N = P.N, pl = P.nx * P.ny;
    const float phi_C = fld[c];
    const float phi_E = (i != P.nx - 1) ? fld[c + 1] : 0.0f;
    const float phi_W = (i != 0) ? fld[c - 1] : 0.0f;
    const float phi_N = (j != P.ny - 1) ? fld[c + P.nx] : 0.0f;
    const float phi_S = (j != 0) ? fld[c - P.nx] : 0.0f;
    float newPhi = src[c] - coef[c + N * sE] * phi_E - coef[c + N * sW] * phi_W - coef[c + N * sN] * phi_N
                   - coef[c + N * sS] * phi_S;
    if (DIM == 3) {
        const float phi_T = (k != P.nz - 1) ? fld[c + pl] : 0.0f;
        const float phi_B = (k != 0) ? fld[c - pl] : 0.0f;
        newPhi -= coef[c + N * sT] * phi_T + coef[c + N * sB] * phi_B;
    }
    newPhi /= coef[c + N * sC];
    dPhi = rlx * (newPhi - phi_C);
    return phi_C + dPhi;
}

// ---------------- point Jacobi sweep, kernels.cuh:845-924 ----------------
// Reads `in`, writes `out` (ping-pong; a sweep that does not run copies), so after an even number
// of launches the result is in the caller's field and an odd count ends with a copy back:
// the reference's pointer swap (Solver.cu:894-896) with the result always in the caller's field.
template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_sweep_jacobi(CaseParams P0, BatchMap M, SweepSel e, int launch) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    const bool run = e.k > 0 && launchRuns(P, blk, e.slot, launch, e.kPrev, blk.n, P.tol[e.q]);
    float *fld = slotField(P, e.slot), *alt = slotAlt(P, e.slot);
    const bool odd = ((launch + e.flip) & 1) != 0;
    const float *in = odd ? alt : fld;
    float *out = odd ? fld : alt;
    const float *coef = slotCoef(P, e.slot), *src = slotSource(P, e.slot);
    const float rlx = P.relax[e.q];
    double s[1] = {0.0};
    forCells(P, blk, [&](int i, int j, int k, int c) {
        if (run) {
            float d;
            out[c] = gsUpdate<DIM>(P, in, coef, src, i, j, k, c, rlx, d);
            s[0] += (double)d * (double)d;
        } else {
            out[c] = in[c];
        }
    });
    if (run) storeBlockSums<1>(s, sweepPart(P, e.slot, launch, 0) + blk.flat);
}

// Closes a smoother call of nL launches (the last one ran kLast sweeps): publishes the state
// after the last launch and adds the sweeps that ran to the outer iteration's count.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_smooth_finish(CaseParams P0, BatchMap M, int slot, int q, int nL, int kLast,
                                                          int sweepKind, int nprev) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_ONE, blk);
    if (caseDone(P)) return;
    launchRuns(P, blk, slot, nL, kLast, kindCount<BATCH>(M, sweepKind, P, nprev), P.tol[q]);
    if (threadIdx.x == 0) P.st->sweepsIter[q] += P.st->sm[slot].sweeps;
}

// dir 0: a sweep call that ended after an odd number of launches left its result in the partner
// buffer; copy it back.  dir 1: copy the field into its partner (keeps them equal).
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_copy_back(CaseParams P0, BatchMap M, int slot, int dir) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_COPY, blk);
    if (caseDone(P)) return;
    const int t = blk.flat * kBlock + threadIdx.x;
    if (t < P.N) {
        if (dir == 0) slotField(P, slot)[t] = slotAlt(P, slot)[t];
        else slotAlt(P, slot)[t] = slotField(P, slot)[t];
    }
}

// ---------------- Rhie-Chow, kernels.cuh:336-500, then face BCs (Solver.cu:187) ----------------
__device__ __forceinline__ float rhieChow(const CaseParams &P, int d, int c, int idx, int n, int st,
                                          const float *vel, const float *aCp) {
    const int lo = 2 * d + 1, hi = 2 * d;
    const float area = d == 0 ? P.Ax : d == 1 ? P.Ay : P.Az;
    const int LId = c - st, RId = c;
    const float aC_L = aCp[LId];
    const float aC_R = aCp[RId];
    const float u_L = vel[LId];
    const float u_R = vel[RId];
    const float p_L = P.p[LId], p_R = P.p[RId];
    float p_LL, p_RR;
    if (idx == 1) {
        p_LL = ghostCopy(P, lo) ? p_L : 0.0f;
        p_RR = P.p[RId + st];
    } else if (idx == n - 1) {
        p_RR = ghostCopy(P, hi) ? p_R : 0.0f;
        p_LL = P.p[LId - st];
    } else {
        p_LL = P.p[LId - st];
        p_RR = P.p[RId + st];
    }
    return 0.5 * (u_L + u_R) + (p_R - p_LL) * area / (4 * aC_L) + (p_RR - p_L) * area / (4 * aC_R)
           + 0.5 * (1 / aC_L + 1 / aC_R) * (p_L - p_R) * area;
}

template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_rhie_chow(CaseParams P0, BatchMap M, Finish fin) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    runFinish<BATCH>(P, M, blk, fin);
    const float *aC = P.mcoef;   // plane sC
    forCells(P, blk, [&](int i, int j, int k, int c) {
        if (i >= 1) P.uf[ufi(P, i, j, k)] = rhieChow(P, 0, c, i, P.nx, 1, P.u, aC);
        if (j >= 1) P.vf[vfi(P, i, j, k)] = rhieChow(P, 1, c, j, P.ny, P.nx, P.v, aC);
        if (DIM == 3 && k >= 1) P.wf[wfi(P, i, j, k)] = rhieChow(P, 2, c, k, P.nz, P.nx * P.ny, P.w, aC);
        applyFaceBCs<DIM>(P, i, j, k, c);
    });
}

// ---------------- pressure-correction assembly, kernels.cuh:502-581, and p' = 0 ----------------
template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_pcorr_assemble(CaseParams P0, BatchMap M) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    double s[1] = {0.0};
    forCells(P, blk, [&](int i, int j, int k, int c) {
        const int N = P.N, pl = P.nx * P.ny;
        const float *uC = P.mcoef;
        const float rho = P.rho, Ax = P.Ax, Ay = P.Ay, Az = P.Az;
        float aE = 0.0f, aW = 0.0f, aN = 0.0f, aS = 0.0f, aT = 0.0f, aB = 0.0f;
        if (i < P.nx - 1) aE = -0.5 * rho * Ax * Ax * (1 / uC[c] + 1 / uC[c + 1]);
        if (i > 0) aW = -0.5 * rho * Ax * Ax * (1 / uC[c] + 1 / uC[c - 1]);
        if (j < P.ny - 1) aN = -0.5 * rho * Ay * Ay * (1 / uC[c] + 1 / uC[c + P.nx]);
        if (j > 0) aS = -0.5 * rho * Ay * Ay * (1 / uC[c] + 1 / uC[c - P.nx]);
        float aCc = -(aE + aW + aN + aS);
        if (DIM == 3) {
            if (k < P.nz - 1) aT = -0.5 * rho * Az * Az * (1 / uC[c] + 1 / uC[c + pl]);
            if (k > 0) aB = -0.5 * rho * Az * Az * (1 / uC[c] + 1 / uC[c - pl]);
            aCc += -(aT + aB);
        }
        float *pcf = P.pcoef;
        P.rinv[c] = 1 / uC[c];
        pcf[c] = aCc;
        pcf[c + N * sE] = aE;
        pcf[c + N * sW] = aW;
        pcf[c + N * sN] = aN;
        pcf[c + N * sS] = aS;
        if (DIM == 3) {
            pcf[c + N * sT] = aT;
            pcf[c + N * sB] = aB;
        }
        float b = rho * (Ax * (P.uf[ufi(P, i, j, k)] - P.uf[ufi(P, i + 1, j, k)])
                         + Ay * (P.vf[vfi(P, i, j, k)] - P.vf[vfi(P, i, j + 1, k)]));
        if (DIM == 3) b += rho * Az * (P.wf[wfi(P, i, j, k)] - P.wf[wfi(P, i, j, k + 1)]);
        P.bp[c] = b;
        P.pc[c] = 0.0f;                       // Solver.cu:193
        P.alt[3][c] = 0.0f;                   // the partner buffer too (odd launch counts start there)
        s[0] += (double)b * (double)b;
    });
    storeBlockSums<1>(s, P.partMass + blk.flat);
}

// ---------------- FP64 defect of the p' system (refinement), delta = 0 ----------------
template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_pcorr_defect(CaseParams P0, BatchMap M, Finish fin) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    runFinish<BATCH>(P, M, blk, fin);
    forCells(P, blk, [&](int i, int j, int k, int c) {
        float a[7];
#pragma unroll
        for (int s = 0; s < 7; ++s) a[s] = (s < 1 + 2 * DIM) ? P.pcoef[c + P.N * s] : 0.0f;
        P.tmp[c] = (float)momResidual<DIM>(P, P.pc, a, P.bp[c], i, j, k, c);
        P.tmp2[c] = 0.0f;
        P.alt[4][c] = 0.0f;
    });
}

template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_pcorr_accumulate(CaseParams P0, BatchMap M) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    forCells(P, blk, [&](int, int, int, int c) { P.pc[c] += P.tmp2[c]; });
}

// ---------------- updateField, kernels.cuh:583-843 ----------------
__device__ __forceinline__ float velCorr(const CaseParams &P, int d, int idx, int n, int c, int st) {
    const int lo = 2 * d + 1, hi = 2 * d;
    const float area = d == 0 ? P.Ax : d == 1 ? P.Ay : P.Az;
    float pc_L, pc_R;
    if (idx == 0) {
        pc_L = ghostCopy(P, lo) ? P.pc[c] : 0.0f;
        pc_R = P.pc[c + st];
    } else if (idx == n - 1) {
        pc_R = ghostCopy(P, hi) ? P.pc[c] : 0.0f;
        pc_L = P.pc[c - st];
    } else {
        pc_L = P.pc[c - st];
        pc_R = P.pc[c + st];
    }
    const float aC_C = P.mcoef[c];
    return P.relax[d] * 0.5 * (pc_L - pc_R) * area / aC_C;
}
__device__ __forceinline__ float faceCorr(const CaseParams &P, int d, int c, int st) {
    const float area = d == 0 ? P.Ax : d == 1 ? P.Ay : P.Az;
    const float aC_L = P.mcoef[c - st];
    const float aC_R = P.mcoef[c];
    return P.relax[d] * 0.5 * (1 / aC_L + 1 / aC_R) * (P.pc[c - st] - P.pc[c]) * area;
}

template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_update(CaseParams P0, BatchMap M, Finish fin) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    runFinish<BATCH>(P, M, blk, fin);
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    forCells(P, blk, [&](int i, int j, int k, int c) {
        const int pl = P.nx * P.ny;
        const float dp = P.relax[3] * P.pc[c];          // kernels.cuh:825
        P.p[c] += dp;
        s[0] += (double)dp * dp;
        // corrected velocities also go to the smoothers' partner buffers, which a momentum call
        // with an odd launch count reads first
        const float du = velCorr(P, 0, i, P.nx, c, 1);
        const float un = P.u[c] + du;
        P.u[c] = un;
        P.alt[0][c] = un;
        s[1] += (double)du * du;
        const float dv = velCorr(P, 1, j, P.ny, c, P.nx);
        const float vn = P.v[c] + dv;
        P.v[c] = vn;
        P.alt[1][c] = vn;
        s[2] += (double)dv * dv;
        if (DIM == 3) {
            const float dw = velCorr(P, 2, k, P.nz, c, pl);
            const float wn = P.w[c] + dw;
            P.w[c] = wn;
            P.alt[2][c] = wn;
            s[3] += (double)dw * dw;
        }
        if (i >= 1) {
            const float d = faceCorr(P, 0, c, 1);
            P.uf[ufi(P, i, j, k)] += d;
            s[4] += (double)d * d;
        }
        if (j >= 1) {
            const float d = faceCorr(P, 1, c, P.nx);
            P.vf[vfi(P, i, j, k)] += d;
            s[5] += (double)d * d;
        }
        if (DIM == 3 && k >= 1) {
            const float d = faceCorr(P, 2, c, pl);
            P.wf[wfi(P, i, j, k)] += d;
            s[6] += (double)d * d;
        }
    });
    storeBlockSums<7>(s, P.partUpd + 7 * blk.flat);
}

// ---------------- outer-iteration monitor: Solver.cu:226-268 (FP64) + scaled residual ----------------
__device__ void finishOuter(const CaseParams &P) {
    CaseState &S = *P.st;
    const double *u = S.upd;
    const double Nd = (double)P.N;
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    ifl_iter_record rec;
    rec.p_norm = sqrt(u[0] / Nd);
    rec.u_norm = sqrt(u[1] / Nd);
    rec.v_norm = sqrt(u[2] / Nd);
    rec.w_norm = 0.0;
    rec.uf_norm = sqrt(u[4] / ((double)(nx - 1) * ny * nz));
    rec.vf_norm = sqrt(u[5] / ((double)nx * (ny - 1) * nz));
    rec.wf_norm = 0.0;
    rec.mass_norm = sqrt(S.mass / Nd);
    if (P.dim == 3) {
        rec.w_norm = sqrt(u[3] / Nd);
        rec.wf_norm = sqrt(u[6] / ((double)nx * ny * (nz - 1)));
    }
    rec.res_cont = sqrt(S.mass);
    rec.res_u = sqrt(S.resSq[0]);
    rec.res_v = sqrt(S.resSq[1]);
    rec.res_w = sqrt(S.resSq[2]);
    rec.res_t = sqrt(S.resSq[3]);
    const double rs[5] = {rec.res_cont, rec.res_u, rec.res_v, rec.res_w, rec.res_t};
    if (S.iter < 5)
        for (int q = 0; q < 5; ++q) S.resMax[q] = fmax(S.resMax[q], rs[q]);
    double sc = 0.0;
    for (int q = 0; q < 5; ++q)
        if ((q < 3 || (q == 3 && P.dim == 3) || (q == 4 && P.energy)) && S.resMax[q] > 0)
            sc = fmax(sc, rs[q] / S.resMax[q]);
    rec.scaled = sc;
    S.pNormMax = fmax(S.pNormMax, rec.p_norm);
    if (P.stopRule == IFL_STOP_FAITHFUL) {
        double m = fmax(S.pNormMax, rec.p_norm);
        S.uNormMax = S.vNormMax = S.ufNormMax = S.vfNormMax = S.wNormMax = S.wfNormMax = m;
    } else {
        S.uNormMax = fmax(S.uNormMax, rec.u_norm);
        S.vNormMax = fmax(S.vNormMax, rec.v_norm);
        S.wNormMax = fmax(S.wNormMax, rec.w_norm);
        S.ufNormMax = fmax(S.ufNormMax, rec.uf_norm);
        S.vfNormMax = fmax(S.vfNormMax, rec.vf_norm);
        S.wfNormMax = fmax(S.wfNormMax, rec.wf_norm);
    }
    rec.rel_p = rec.p_norm / S.pNormMax;
    rec.rel_u = rec.u_norm / S.uNormMax;
    rec.rel_v = rec.v_norm / S.vNormMax;
    rec.rel_uf = rec.uf_norm / S.ufNormMax;
    rec.rel_vf = rec.vf_norm / S.vfNormMax;
    rec.rel_w = 0.0;
    rec.rel_wf = 0.0;
    if (P.dim == 3) {
        rec.rel_w = rec.w_norm / S.wNormMax;
        rec.rel_wf = rec.wf_norm / S.wfNormMax;
    }
    bool conv = false;
    if (P.stopRule == IFL_STOP_FAITHFUL || P.stopRule == IFL_STOP_FIXED) {
        conv = rec.rel_p < P.tol[3] && rec.rel_u < P.tol[0] && rec.rel_v < P.tol[1] && rec.rel_uf < P.tol[0]
               && rec.rel_vf < P.tol[1];
        if (P.dim == 3) conv = conv && rec.rel_w < P.tol[2] && rec.rel_wf < P.tol[2];
        conv = conv || rec.mass_norm < P.tolMass;
    } else if (P.stopRule == IFL_STOP_SCALED) {
        conv = sc < P.tolRes;
    }
    for (int q = 0; q < 4; ++q) rec.sweeps[q] = S.sweepsIter[q];
    rec.sweeps_t = S.sweepsIter[4];
    rec.converged = conv ? 1 : 0;
    S.lastScaled = sc;
    S.lastConverged = rec.converged;
    if (P.hist) P.hist[S.iter % P.histCap] = rec;
    const double chk[] = {rec.p_norm, rec.u_norm, rec.v_norm, rec.w_norm, rec.uf_norm, rec.vf_norm,
                          rec.wf_norm, rec.mass_norm, rec.res_cont, rec.res_u, rec.res_v, rec.res_w, rec.res_t};
    bool finite = true;
    for (double x : chk) finite = finite && isfinite(x);
    S.iter += 1;
    if (!finite) S.done = 2;
    else if (conv) S.done = 1;
    else if (S.iterCap > 0 && S.iter >= S.iterCap) S.done = 3;
}

// One workgroup: totals of the cell-kernel partials (fixed order), then (optionally) the monitor.
// All three partial arrays are read in one pass, four records per thread in flight at a time.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_reduce_cells(CaseParams P0, BatchMap M, int nSingle, int mask) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_ONE, blk);
    if (caseDone(P)) return;
    const int nblk = kindCount<BATCH>(M, MAP_CELL, P, nSingle);
    const bool dm = (mask & RED_MOM) != 0, dM = (mask & RED_MASS) != 0, du = (mask & RED_UPD) != 0;
    double acc[12];
#pragma unroll
    for (int s = 0; s < 12; ++s) acc[s] = 0.0;
    // record b adds in the order of b (thread t owns b = t, t + 256, ...; chunks of four keep the order)
    for (int b0 = threadIdx.x; b0 < nblk; b0 += 4 * kBlock) {
        double r[4][12];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b = b0 + q * kBlock;
            const bool ok = b < nblk;
#pragma unroll
            for (int s = 0; s < 4; ++s) r[q][s] = (ok && dm) ? P.partMom[4 * b + s] : 0.0;
            r[q][4] = (ok && dM) ? P.partMass[b] : 0.0;
#pragma unroll
            for (int s = 0; s < 7; ++s) r[q][5 + s] = (ok && du) ? P.partUpd[7 * b + s] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int s = 0; s < 12; ++s) acc[s] += r[q][s];
    }
    double tot[12];
    blockSum<12>(acc, tot);
    if (threadIdx.x != 0) return;
    CaseState &S = *P.st;
    if (dm)
        for (int q = 0; q < 4; ++q) S.resSq[q] = tot[q];
    if (dM) S.mass = tot[4];
    if (du)
        for (int q = 0; q < 7; ++q) S.upd[q] = tot[5 + q];
    if (mask & RED_FINISH) finishOuter(P);
}

// ---------------- host-side launchers ----------------
dim3 cellGrid(const CaseParams &P) {
    return dim3((unsigned)((P.nx + 63) / 64), (unsigned)((P.ny * P.nz + 4 * kRows - 1) / (4 * kRows)));
}
int cell_blocks(const CaseParams &P) {
    dim3 g = cellGrid(P);
    return (int)(g.x * g.y);
}
int cell_dimx(const CaseParams &P) { return (P.nx + 63) / 64; }
int copy_blocks(const CaseParams &P) { return (P.N + kBlock - 1) / kBlock; }

static dim3 gridOf(const Lc &L, int kind) {
    if (L.batch) return dim3((unsigned)L.total[kind]);
    switch (kind) {
        case MAP_CELL: return cellGrid(*L.P);
        case MAP_ONE: return dim3(1);
        case MAP_COPY: return dim3((unsigned)copy_blocks(*L.P));
        default: return dim3(1);
    }
}

#define IFL_LAUNCH(L, KIND, KERNEL, ...)                                                               \
    do {                                                                                              \
        if ((L).batch) KERNEL<true><<<gridOf((L), (KIND)), kBlock, 0, (L).s>>>(*(L).P, (L).M, ##__VA_ARGS__);  \
        else KERNEL<false><<<gridOf((L), (KIND)), kBlock, 0, (L).s>>>(*(L).P, (L).M, ##__VA_ARGS__);          \
    } while (0)
#define IFL_LAUNCH_DIM(L, KIND, KERNEL, ...)                                                           \
    do {                                                                                              \
        const int d_ = (L).P->dim;                                                                    \
        if ((L).batch) {                                                                              \
            if (d_ == 3) KERNEL<3, true><<<gridOf((L), (KIND)), kBlock, 0, (L).s>>>(*(L).P, (L).M, ##__VA_ARGS__); \
            else KERNEL<2, true><<<gridOf((L), (KIND)), kBlock, 0, (L).s>>>(*(L).P, (L).M, ##__VA_ARGS__);         \
        } else {                                                                                      \
            if (d_ == 3) KERNEL<3, false><<<gridOf((L), (KIND)), kBlock, 0, (L).s>>>(*(L).P, (L).M, ##__VA_ARGS__); \
            else KERNEL<2, false><<<gridOf((L), (KIND)), kBlock, 0, (L).s>>>(*(L).P, (L).M, ##__VA_ARGS__);         \
        }                                                                                             \
    } while (0)

void launch_init_faces(const Lc &L) { IFL_LAUNCH_DIM(L, MAP_CELL, k_init_faces); }
void launch_face_bc(const Lc &L) { IFL_LAUNCH_DIM(L, MAP_CELL, k_face_bc); }
void launch_mom_assemble(const Lc &L) { IFL_LAUNCH_DIM(L, MAP_CELL, k_mom_assemble); }
void launch_rhie_chow(const Lc &L, const Finish &fin) { IFL_LAUNCH_DIM(L, MAP_CELL, k_rhie_chow, fin); }
void launch_pcorr_assemble(const Lc &L) { IFL_LAUNCH_DIM(L, MAP_CELL, k_pcorr_assemble); }
void launch_update(const Lc &L, const Finish &fin) { IFL_LAUNCH_DIM(L, MAP_CELL, k_update, fin); }
void launch_pcorr_defect(const Lc &L, const Finish &fin) { IFL_LAUNCH_DIM(L, MAP_CELL, k_pcorr_defect, fin); }
void launch_pcorr_accumulate(const Lc &L) { IFL_LAUNCH(L, MAP_CELL, k_pcorr_accumulate); }
void launch_reduce_cells(const Lc &L, int mask) { IFL_LAUNCH(L, MAP_ONE, k_reduce_cells, cell_blocks(*L.P), mask); }
void launch_sweep_jacobi(const Lc &L, const SweepSel &e, int launch) {
    IFL_LAUNCH_DIM(L, MAP_CELL, k_sweep_jacobi, e, launch);
}
void launch_smooth_finish(const Lc &L, int slot, int q, int nL, int kLast, int sweepKind, int nprev) {
    IFL_LAUNCH(L, MAP_ONE, k_smooth_finish, slot, q, nL, kLast, sweepKind, nprev);
}
void launch_copy_back(const Lc &L, int slot, int dir) { IFL_LAUNCH(L, MAP_COPY, k_copy_back, slot, dir); }

}  // namespace ifl
