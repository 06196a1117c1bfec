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
    const int R = P.cellRows;
    const int r0 = b.y * (4 * R) + (threadIdx.x >> 6);
#pragma unroll
    for (int q = 0; q < kRows; ++q) {
        const int r = r0 + 4 * q;
        if (q < R && i < P.nx && r < rows) {
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
    const LinkConsts lkU = linkConsts<DIM>(P, P.mu), lkT = linkConsts<DIM>(P, P.gam);
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
        linkCoef<DIM>(lkU, ue, uw, vn, vs, wt, wb, a);
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
            linkCoef<DIM>(lkT, ue, uw, vn, vs, wt, wb, t);
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
// The face between cells L = c - st and R = c of direction d (idx = R's index along d, 1..n-1),
// from the loaded pressures p_LL, p_L, p_R, p_RR (p_LL read at c when idx == 1, p_RR at c when
// idx == n - 1: the ghost rules of kernels.cuh:481-495 replace them), velocities and diagonals.
__device__ __forceinline__ float rhieChow(const CaseParams &P, int d, int idx, int n, float u_L, float u_R,
                                          float aC_L, float aC_R, float p_LLr, float p_L, float p_R, float p_RRr) {
    const int lo = 2 * d + 1, hi = 2 * d;
    const float area = d == 0 ? P.Ax : d == 1 ? P.Ay : P.Az;
    const float p_LL = (idx == 1) ? (ghostCopy(P, lo) ? p_L : 0.0f) : p_LLr;
    const float p_RR = (idx == n - 1) ? (ghostCopy(P, hi) ? p_R : 0.0f) : p_RRr;
    return 0.5 * (u_L + u_R) + (p_R - p_LL) * area / (4 * aC_L) + (p_RR - p_L) * area / (4 * aC_R)
           + 0.5 * (1 / aC_L + 1 / aC_R) * (p_L - p_R) * area;
}

// Profile values of the faces of cell (i,j,k) (read from u where a face has none; then unused).
template <int DIM>
__device__ __forceinline__ void loadProfiles(const CaseParams &P, int i, int j, int k, float (&prv)[6]) {
    const int nx = P.nx, ny = P.ny;
    const int qf[6] = {j + ny * k, j + ny * k, i + nx * k, i + nx * k, i + nx * j, i + nx * j};
#pragma unroll
    for (int f = 0; f < 6; ++f) prv[f] = f < 2 * DIM ? (P.prof[f] != nullptr ? P.prof[f] : P.u)[qf[f]] : 0.0f;
}

// applyFaceBCs with the cell's velocities and face profile values already loaded.
template <int DIM>
__device__ __forceinline__ void storeFaceBCs(const CaseParams &P, int i, int j, int k, float uc, float vc, float wc,
                                             const float (&prv)[6]) {
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    if (i == 0) P.uf[ufi(P, 0, j, k)] = faceBCv(P, W, prv[W], uc);
    if (i == nx - 1) P.uf[ufi(P, nx, j, k)] = faceBCv(P, E, prv[E], uc);
    if (j == 0) P.vf[vfi(P, i, 0, k)] = faceBCv(P, S, prv[S], vc);
    if (j == ny - 1) P.vf[vfi(P, i, ny, k)] = faceBCv(P, N, prv[N], vc);
    if (DIM == 3) {
        if (k == 0) P.wf[wfi(P, i, j, 0)] = faceBCv(P, B, prv[B], wc);
        if (k == nz - 1) P.wf[wfi(P, i, j, nz)] = faceBCv(P, T, prv[T], wc);
    }
}

template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_rhie_chow(CaseParams P0, BatchMap M, Finish fin) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    runFinish<BATCH>(P, M, blk, fin);
    const float *aC = P.mcoef;   // plane sC
    forCells(P, blk, [&](int i, int j, int k, int c) {
        const int n3[3] = {P.nx, P.ny, P.nz}, id3[3] = {i, j, k};
        const int st3[3] = {1, P.nx, P.nx * P.ny};
        const float *vel3[3] = {P.u, P.v, P.w};
        // every load first: for each direction the cells c-2st, c-st, c, c+st (at c where missing)
        float pLL[3], pL[3], pR[3], pRR[3], uL[3], uR[3], aL[3], aR[3];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int st = st3[d], idx = id3[d], n = n3[d];
            const int cL = idx >= 1 ? c - st : c, cLL = idx >= 2 ? c - 2 * st : c, cRR = idx <= n - 2 ? c + st : c;
            pLL[d] = P.p[cLL];
            pL[d] = P.p[cL];
            pR[d] = P.p[c];
            pRR[d] = P.p[cRR];
            uL[d] = vel3[d][cL];
            uR[d] = vel3[d][c];
            aL[d] = aC[cL];
            aR[d] = aC[c];
        }
        float prv[6];
        loadProfiles<DIM>(P, i, j, k, prv);
        if (i >= 1) P.uf[ufi(P, i, j, k)] = rhieChow(P, 0, i, P.nx, uL[0], uR[0], aL[0], aR[0], pLL[0], pL[0], pR[0], pRR[0]);
        if (j >= 1) P.vf[vfi(P, i, j, k)] = rhieChow(P, 1, j, P.ny, uL[1], uR[1], aL[1], aR[1], pLL[1], pL[1], pR[1], pRR[1]);
        if (DIM == 3 && k >= 1)
            P.wf[wfi(P, i, j, k)] = rhieChow(P, 2, k, P.nz, uL[2], uR[2], aL[2], aR[2], pLL[2], pL[2], pR[2], pRR[2]);
        storeFaceBCs<DIM>(P, i, j, k, uR[0], uR[1], DIM == 3 ? uR[2] : 0.0f, prv);
    });
}

// ---------------- pressure-correction assembly, kernels.cuh:502-581, and p' = 0 ----------------
template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_pcorr_assemble(CaseParams P0, BatchMap M, int writeAlt) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    double s[1] = {0.0};
    forCells(P, blk, [&](int i, int j, int k, int c) {
        const int N = P.N;
        const float rho = P.rho, Ax = P.Ax, Ay = P.Ay, Az = P.Az;
        const Nbr nb = nbrOf<DIM>(P, i, j, k);
        float uC[7];
        gather7<DIM>(P.mcoef, c, nb, uC);          // the momentum diagonal (plane sC)
        const float ufW = P.uf[ufi(P, i, j, k)], ufE = P.uf[ufi(P, i + 1, j, k)];
        const float vfS = P.vf[vfi(P, i, j, k)], vfN = P.vf[vfi(P, i, j + 1, k)];
        const float wfB = DIM == 3 ? P.wf[wfi(P, i, j, k)] : 0.0f, wfT = DIM == 3 ? P.wf[wfi(P, i, j, k + 1)] : 0.0f;
        float aE = 0.0f, aW = 0.0f, aN = 0.0f, aS = 0.0f, aT = 0.0f, aB = 0.0f;
        if (nb.has[1]) aE = -0.5 * rho * Ax * Ax * (1 / uC[0] + 1 / uC[1]);
        if (nb.has[2]) aW = -0.5 * rho * Ax * Ax * (1 / uC[0] + 1 / uC[2]);
        if (nb.has[3]) aN = -0.5 * rho * Ay * Ay * (1 / uC[0] + 1 / uC[3]);
        if (nb.has[4]) aS = -0.5 * rho * Ay * Ay * (1 / uC[0] + 1 / uC[4]);
        float aCc = -(aE + aW + aN + aS);
        if (DIM == 3) {
            if (nb.has[5]) aT = -0.5 * rho * Az * Az * (1 / uC[0] + 1 / uC[5]);
            if (nb.has[6]) aB = -0.5 * rho * Az * Az * (1 / uC[0] + 1 / uC[6]);
            aCc += -(aT + aB);
        }
        float *pcf = P.pcoef;
        P.rinv[c] = 1 / uC[0];
        pcf[c] = aCc;
        pcf[c + N * sE] = aE;
        pcf[c + N * sW] = aW;
        pcf[c + N * sN] = aN;
        pcf[c + N * sS] = aS;
        if (DIM == 3) {
            pcf[c + N * sT] = aT;
            pcf[c + N * sB] = aB;
        }
        float b = rho * (Ax * (ufW - ufE) + Ay * (vfS - vfN));
        if (DIM == 3) b += rho * Az * (wfB - wfT);
        P.bp[c] = b;
        P.pc[c] = 0.0f;                       // Solver.cu:193
        if (writeAlt) P.alt[3][c] = 0.0f;     // the partner too, when the p' call starts there
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
        const Nbr nb = nbrOf<DIM>(P, i, j, k);
        float a[7], x[7];
#pragma unroll
        for (int s = 0; s < 7; ++s) a[s] = (s < 1 + 2 * DIM) ? P.pcoef[c + P.N * s] : 0.0f;
        gather7<DIM>(P.pc, c, nb, x);
        const float b = P.bp[c];
        P.tmp[c] = (float)residual7<DIM>(a, b, x, nb);
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
// updateInteriorVelKernel / updateBCVelKernel (:583-748) on the loaded p' neighbourhood: direction
// d, position idx of n cells.
__device__ __forceinline__ float velCorr7(const CaseParams &P, int d, int idx, int n, const float (&pc)[7], float aC_C) {
    const int lo = 2 * d + 1, hi = 2 * d;
    const float area = d == 0 ? P.Ax : d == 1 ? P.Ay : P.Az;
    const float pc_L = (idx == 0) ? (ghostCopy(P, lo) ? pc[0] : 0.0f) : pc[2 + 2 * d];
    const float pc_R = (idx == n - 1) ? (ghostCopy(P, hi) ? pc[0] : 0.0f) : pc[1 + 2 * d];
    return P.relax[d] * 0.5 * (pc_L - pc_R) * area / aC_C;
}
// updateFaceVelKernel (:750-812) for the face between the cell's lower neighbour along d and the cell.
__device__ __forceinline__ float faceCorr7(const CaseParams &P, int d, float aC_L, float aC_R, float pc_L, float pc_R) {
    const float area = d == 0 ? P.Ax : d == 1 ? P.Ay : P.Az;
    return P.relax[d] * 0.5 * (1 / aC_L + 1 / aC_R) * (pc_L - pc_R) * area;
}

template <int DIM, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_update(CaseParams P0, BatchMap M, Finish fin, int writeAlt) {
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_CELL, blk);
    if (caseDone(P)) return;
    runFinish<BATCH>(P, M, blk, fin);
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    forCells(P, blk, [&](int i, int j, int k, int c) {
        const Nbr nb = nbrOf<DIM>(P, i, j, k);
        // every load first
        float pc[7], aC[7];
        gather7<DIM>(P.pc, c, nb, pc);
        gather7<DIM>(P.mcoef, c, nb, aC);
        const float p0 = P.p[c], u0 = P.u[c], v0 = P.v[c], w0 = DIM == 3 ? P.w[c] : 0.0f;
        const float fu = P.uf[ufi(P, i, j, k)], fv = P.vf[vfi(P, i, j, k)];
        const float fw = DIM == 3 ? P.wf[wfi(P, i, j, k)] : 0.0f;
        const float dp = P.relax[3] * pc[0];            // kernels.cuh:825
        P.p[c] = p0 + dp;
        s[0] += (double)dp * dp;
        // corrected velocities also go to the smoothers' partner buffers when a momentum call
        // with an odd launch count reads them first (writeAlt)
        const float du = velCorr7(P, 0, i, P.nx, pc, aC[0]);
        const float un = u0 + du;
        P.u[c] = un;
        if (writeAlt) P.alt[0][c] = un;
        s[1] += (double)du * du;
        const float dv = velCorr7(P, 1, j, P.ny, pc, aC[0]);
        const float vn = v0 + dv;
        P.v[c] = vn;
        if (writeAlt) P.alt[1][c] = vn;
        s[2] += (double)dv * dv;
        if (DIM == 3) {
            const float dw = velCorr7(P, 2, k, P.nz, pc, aC[0]);
            const float wn = w0 + dw;
            P.w[c] = wn;
            if (writeAlt) P.alt[2][c] = wn;
            s[3] += (double)dw * dw;
        }
        if (i >= 1) {
            const float d = faceCorr7(P, 0, aC[2], aC[0], pc[2], pc[0]);
            P.uf[ufi(P, i, j, k)] = fu + d;
            s[4] += (double)d * d;
        }
        if (j >= 1) {
            const float d = faceCorr7(P, 1, aC[4], aC[0], pc[4], pc[0]);
            P.vf[vfi(P, i, j, k)] = fv + d;
            s[5] += (double)d * d;
        }
        if (DIM == 3 && k >= 1) {
            const float d = faceCorr7(P, 2, aC[6], aC[0], pc[6], pc[0]);
            P.wf[wfi(P, i, j, k)] = fw + d;
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
    blockSumWave0<12>(acc, tot);
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
    return dim3((unsigned)((P.nx + 63) / 64), (unsigned)((P.ny * P.nz + 4 * P.cellRows - 1) / (4 * P.cellRows)));
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
void launch_pcorr_assemble(const Lc &L, bool writeAlt) { IFL_LAUNCH_DIM(L, MAP_CELL, k_pcorr_assemble, writeAlt ? 1 : 0); }
void launch_update(const Lc &L, const Finish &fin, bool writeAlt) {
    IFL_LAUNCH_DIM(L, MAP_CELL, k_update, fin, writeAlt ? 1 : 0);
}
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
