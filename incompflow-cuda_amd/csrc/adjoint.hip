// adjoint.hip - discrete adjoint of the converged SIMPLE state on gfx950.
//
// One thread per cell (the cell grid of kernels.hip).  Each kernel computes, for its cell, the
// adjoint elements that land on that cell as a sum over the cells and faces that reach it, in
// the fixed order of oracle/adjoint_oracle.hpp, so no atomics are needed and the result is
// deterministic.  The reference has no adjoint; the stages transposed are those of one outer
// iteration of Solver.cu:160-290 (kernels.cuh:97-843) at its fixed point.
#include "ifl_adjoint.hpp"
#include "ifl_stencil.hpp"

namespace ifl {

namespace {

template <class F>
__device__ __forceinline__ void forCellsA(const CaseParams &P, F &&f) {
    const int rows = P.ny * P.nz;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r0 = blockIdx.y * (4 * kRows) + (threadIdx.x >> 6);
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

__device__ __forceinline__ int flatBlock() { return blockIdx.x + gridDim.x * blockIdx.y; }

__device__ __forceinline__ int nDir(const CaseParams &P, int d) { return d == 0 ? P.nx : d == 1 ? P.ny : P.nz; }
__device__ __forceinline__ int strideOf(const CaseParams &P, int d) { return d == 0 ? 1 : d == 1 ? P.nx : P.nx * P.ny; }
__device__ __forceinline__ float areaOf(const CaseParams &P, int d) { return d == 0 ? P.Ax : d == 1 ? P.Ay : P.Az; }
__device__ __forceinline__ float spacingOf(const CaseParams &P, int d) { return d == 0 ? P.dx : d == 1 ? P.dy : P.dz; }
__device__ __forceinline__ int idxOf(int d, int i, int j, int k) { return d == 0 ? i : d == 1 ? j : k; }
__device__ __forceinline__ int loSlot(int d) { return d == 0 ? sW : d == 1 ? sS : sB; }
__device__ __forceinline__ int hiSlot(int d) { return d == 0 ? sE : d == 1 ? sN : sT; }
__device__ __forceinline__ int loFace(int d) { return 2 * d + 1; }
__device__ __forceinline__ int hiFace(int d) { return 2 * d; }
// lo face (W/S/B) of the cell m steps along d from (i,j,k)
__device__ __forceinline__ int faceLoAt(const CaseParams &P, int d, int i, int j, int k, int m) {
    return d == 0 ? ufi(P, i + m, j, k) : d == 1 ? vfi(P, i, j + m, k) : wfi(P, i, j, k + m);
}
__device__ __forceinline__ int faceQ(const CaseParams &P, int f, int i, int j, int k) {
    return f < 2 ? j + P.ny * k : f < 4 ? i + P.nx * k : i + P.nx * j;
}
__device__ __forceinline__ const float *faceArr(const CaseParams &P, int d) { return d == 0 ? P.uf : d == 1 ? P.vf : P.wf; }
__device__ __forceinline__ const float *velOf(const CaseParams &P, int d) { return d == 0 ? P.u : d == 1 ? P.v : P.w; }
__device__ __forceinline__ float aCof(const CaseParams &P, int c) { return P.mcoef[c]; }
__device__ __forceinline__ float wOf(const float *g, int c) { return g ? g[c] : 0.0f; }

// Rhie-Chow pressure stencil of the face whose R cell is Rc (index ridx along d)
__device__ __forceinline__ void rcStencil(const CaseParams &P, int d, int Rc, int ridx, float &pL, float &pR, float &pLL,
                                          float &pRR) {
    const int n = nDir(P, d), st = strideOf(P, d), L = Rc - st;
    pL = P.p[L];
    pR = P.p[Rc];
    if (ridx == 1) {
        pLL = ghostCopy(P, loFace(d)) ? pL : 0.0f;
        pRR = P.p[Rc + st];
    } else if (ridx == n - 1) {
        pRR = ghostCopy(P, hiFace(d)) ? pR : 0.0f;
        pLL = P.p[L - st];
    } else {
        pLL = P.p[L - st];
        pRR = P.p[Rc + st];
    }
}

__device__ __forceinline__ void rcWeights(const CaseParams &P, int d, int Rc, float &kL, float &kR, float &m) {
    const int Lc = Rc - strideOf(P, d);
    const float A = areaOf(P, d);
    kL = A / (4 * aCof(P, Lc));
    kR = A / (4 * aCof(P, Rc));
    m = 0.5f * (1 / aCof(P, Lc) + 1 / aCof(P, Rc)) * A;
}

// u*-bar on the face between Lc and Rc: lambda face plus the divergence adjoint
__device__ __forceinline__ float faceStarBar(const CaseParams &P, const AdjBufs &A, int d, int Lc, int Rc, int fi) {
    return A.lf[d][fi] + P.rho * areaOf(P, d) * (A.bpc[Rc] - A.bpc[Lc]);
}

}  // namespace

// ---------------- T5: p'-bar from the corrections ----------------
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_adj_pcorr(CaseParams P, AdjBufs A) {
    forCellsA(P, [&](int i, int j, int k, int c) {
        float acc = P.relax[3] * A.lp[c];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(P, d), st = strideOf(P, d);
            const float Ar = areaOf(P, d), rl = P.relax[d];
            const float *lud = A.lu[d];
            auto gcell = [&](int e) { return rl * 0.5f * Ar / aCof(P, e) * lud[e]; };
            if (idx <= n - 2) acc += gcell(c + st);
            if (idx >= 1) acc -= gcell(c - st);
            if (idx == 0 && ghostCopy(P, loFace(d))) acc += gcell(c);
            if (idx == n - 1 && ghostCopy(P, hiFace(d))) acc -= gcell(c);
            auto hface = [&](int Lc, int Rc, int fi) {
                return rl * 0.5f * (1 / aCof(P, Lc) + 1 / aCof(P, Rc)) * Ar * A.lf[d][fi];
            };
            if (idx <= n - 2) acc += hface(c, c + st, faceLoAt(P, d, i, j, k, 1));
            if (idx >= 1) acc -= hface(c - st, c, faceLoAt(P, d, i, j, k, 0));
        }
        A.pcb[c] = acc;
        A.bpc[c] = 0.0f;
    });
}

// ---------------- one transposed red-black half sweep of colour col ----------------
// col cells: S = relax*phb/aC is added to the source adjoint, phb keeps phb - relax*phb;
// other cells subtract each neighbour's S times the neighbour's coefficient towards them.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_adj_sweep_t(CaseParams P, const float *in, float *out, float *bb,
                                                        const float *coef, float rlx, int col) {
    forCellsA(P, [&](int i, int j, int k, int c) {
        const int N = P.N;
        const float ph = in[c];
        if (((i + j + k) & 1) == col) {
            const float S = rlx * ph / coef[c];
            bb[c] += S;
            out[c] = ph - rlx * ph;
        } else {
            float acc = ph;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const int idx = idxOf(d, i, j, k), n = nDir(P, d), st = strideOf(P, d);
                if (idx != n - 1) {
                    const int e = c + st;
                    acc -= coef[e + N * loSlot(d)] * (rlx * in[e] / coef[e]);
                }
                if (idx != 0) {
                    const int e = c - st;
                    acc -= coef[e + N * hiSlot(d)] * (rlx * in[e] / coef[e]);
                }
            }
            out[c] = acc;
        }
    });
}

// ---------------- T3 (+ divergence of T4): u*-bar, p-bar partial, aC-bar, theta-bar ----------------
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_adj_rc(CaseParams P, AdjBufs A) {
    forCellsA(P, [&](int i, int j, int k, int c) {
        // u*-bar per component
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(P, d), st = strideOf(P, d);
            float acc = A.lu[d][c];
            if (idx >= 1) acc += 0.5f * faceStarBar(P, A, d, c - st, c, faceLoAt(P, d, i, j, k, 0));
            if (idx <= n - 2) acc += 0.5f * faceStarBar(P, A, d, c, c + st, faceLoAt(P, d, i, j, k, 1));
            const float rA = P.rho * areaOf(P, d);
            if (idx == 0 && P.bct[loFace(d)] == IFL_OUTLET) acc += rA * A.bpc[c];
            if (idx == n - 1 && P.bct[hiFace(d)] == IFL_OUTLET) acc += -(rA * A.bpc[c]);
            A.us[d][c] = acc;
            A.rb[d][c] = 0.0f;
        }
        // p-bar: lambda_p + Rhie-Chow pressure terms
        float pacc = A.lp[c];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(P, d), st = strideOf(P, d);
            float kL, kR, m;
            auto w = [&](int mstep) {   // face star-bar of the face whose R cell is mstep steps along d
                const int Rc = c + mstep * st;
                return faceStarBar(P, A, d, Rc - st, Rc, faceLoAt(P, d, i, j, k, mstep));
            };
            if (idx >= 1) {
                rcWeights(P, d, c, kL, kR, m);
                pacc += (kL - m) * w(0);
            }
            if (idx <= n - 2) {
                rcWeights(P, d, c + st, kL, kR, m);
                pacc += (m - kR) * w(1);
            }
            if (idx + 2 <= n - 1) {
                rcWeights(P, d, c + 2 * st, kL, kR, m);
                pacc -= kL * w(2);
            }
            if (idx - 1 >= 1) {
                rcWeights(P, d, c - st, kL, kR, m);
                pacc += kR * w(-1);
            }
            if (idx == 0 && ghostCopy(P, loFace(d))) {
                rcWeights(P, d, c + st, kL, kR, m);
                pacc -= kL * w(1);
            }
            if (idx == n - 1 && ghostCopy(P, hiFace(d))) {
                rcWeights(P, d, c, kL, kR, m);
                pacc += kR * w(0);
            }
        }
        A.pacc[c] = pacc;
        // aC-bar
        float ac = 0.0f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(P, d), st = strideOf(P, d);
            const float Ar = areaOf(P, d), a = aCof(P, c);
            float pL, pR, pLL, pRR;
            if (idx <= n - 2) {
                rcStencil(P, d, c + st, idx + 1, pL, pR, pLL, pRR);
                const float w = faceStarBar(P, A, d, c, c + st, faceLoAt(P, d, i, j, k, 1));
                ac += w * (-((pR - pLL) * Ar / (4 * a * a)) - 0.5f * (pL - pR) * Ar / (a * a));
            }
            if (idx >= 1) {
                rcStencil(P, d, c, idx, pL, pR, pLL, pRR);
                const float w = faceStarBar(P, A, d, c - st, c, faceLoAt(P, d, i, j, k, 0));
                ac += w * (-((pRR - pL) * Ar / (4 * a * a)) - 0.5f * (pL - pR) * Ar / (a * a));
            }
        }
        A.acb[c] = ac;
        // theta-bar of inlet faces through the divergence (first contribution: assignment)
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(P, d);
            const float rA = P.rho * areaOf(P, d);
            if (idx == 0 && P.bct[loFace(d)] == IFL_INLET)
                A.gprof[loFace(d)][faceQ(P, loFace(d), i, j, k)] = (double)(rA * A.bpc[c]);
            if (idx == n - 1 && P.bct[hiFace(d)] == IFL_INLET)
                A.gprof[hiFace(d)][faceQ(P, hiFace(d), i, j, k)] = (double)(-(rA * A.bpc[c]));
        }
    });
}

// ---------------- T1: coefficient adjoints, new p-bar, mu-bar ----------------
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_adj_coef(CaseParams P, AdjBufs A) {
    double red[2] = {0.0, 0.0};   // mu-bar, squared p-bar increment
    forCellsA(P, [&](int i, int j, int k, int c) {
        const int N = P.N;
        float ab[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const float *ud = velOf(P, d);
            const float r = A.rb[d][c];
            ab[sC] -= r * ud[c];
#pragma unroll
            for (int e = 0; e < DIM; ++e) {
                const int idx = idxOf(e, i, j, k), n = nDir(P, e), st = strideOf(P, e);
                if (idx != n - 1) ab[hiSlot(e)] -= r * ud[c + st];
                if (idx != 0) ab[loSlot(e)] -= r * ud[c - st];
            }
        }
        ab[sC] += A.acb[c];
        // BC rows transposed
        bool anyBc = false;
        float pre[7];
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f) {
            const int d = f >> 1;
            const int idx = idxOf(d, i, j, k), n = nDir(P, d);
            const bool on = (f == hiFace(d)) ? idx == n - 1 : idx == 0;
            if (!on) continue;
            if (!anyBc) {
                // coefficients before the BC rows (for the 8/3 source term of inlet faces)
                const float ue = P.uf[ufi(P, i + 1, j, k)], uw = P.uf[ufi(P, i, j, k)];
                const float vn = P.vf[vfi(P, i, j + 1, k)], vs = P.vf[vfi(P, i, j, k)];
                const float wt = DIM == 3 ? P.wf[wfi(P, i, j, k + 1)] : 0.0f;
                const float wb = DIM == 3 ? P.wf[wfi(P, i, j, k)] : 0.0f;
                momLinkCoef<DIM>(P, ue, uw, vn, vs, wt, wb, pre);
                anyBc = true;
            }
            const int q = faceQ(P, f, i, j, k);
            const int ex = bcExSlot(f), in = bcInSlot(f);
            if (P.bct[f] == IFL_WALL || P.bct[f] == IFL_INLET) {
                float a = -2 * ab[sC] + ab[in] / 3;
#pragma unroll
                for (int comp = 0; comp < DIM; ++comp) a -= 8.0f / 3 * bcval(P, f, comp, q) * A.rb[comp][c];
                ab[ex] = a;
                if (P.bct[f] == IFL_INLET) A.gprof[f][q] += (double)(-(8.0f / 3 * pre[ex] * A.rb[d][c]));
            } else {
                ab[ex] = ab[sC];
            }
        }
        A.abar[c] = ab[sC];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            A.abar[c + N * hiSlot(d)] = ab[hiSlot(d)];
            A.abar[c + N * loSlot(d)] = ab[loSlot(d)];
        }
        // mu-bar part of this cell
        double m = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double w = (double)areaOf(P, d) / (double)spacingOf(P, d);
            m += w * ((double)ab[sC] - (double)ab[hiSlot(d)]);
            m += w * ((double)ab[sC] - (double)ab[loSlot(d)]);
        }
        red[0] += m;
        // new p-bar: Rhie-Chow part + pressure sources of the momentum equations + objective
        float ps = 0.0f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(P, d), st = strideOf(P, d);
            const float Ar = areaOf(P, d);
            const float *r = A.rb[d];
            if (idx <= n - 2) ps += 0.5f * Ar * r[c + st];
            if (idx >= 1) ps -= 0.5f * Ar * r[c - st];
            if (idx == 0 && ghostCopy(P, loFace(d))) ps += 0.5f * Ar * r[c];
            if (idx == n - 1 && ghostCopy(P, hiFace(d))) ps -= 0.5f * Ar * r[c];
        }
        const float np = A.pacc[c] + ps + wOf(A.gp, c);
        const double e = (double)np - (double)A.lp[c];
        red[1] += e * e;
        A.lp[c] = np;
    });
    storeBlockSums<2>(red, A.part + 2 * flatBlock());
}

// ---------------- T0/T1 faces: new u-bar and face adjoints, theta-bar ----------------
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_adj_face(CaseParams P, AdjBufs A) {
    double red[6] = {0, 0, 0, 0, 0, 0};   // squared increments u, v, w, uf, vf, wf
    forCellsA(P, [&](int i, int j, int k, int c) {
        const int N = P.N;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(P, d), st = strideOf(P, d);
            const float rA = P.rho * areaOf(P, d);
            const float *uf = faceArr(P, d);
            const int lo = loSlot(d), hi = hiSlot(d);
            float o = A.us[d][c];
            if (idx == 0) {
                const float ub = uf[faceLoAt(P, d, i, j, k, 0)];
                const float g = ub > 0 ? rA * (A.abar[c] - A.abar[c + N * lo]) : 0.0f;
                if (P.bct[loFace(d)] == IFL_OUTLET) o += g;
                else if (P.bct[loFace(d)] == IFL_INLET) A.gprof[loFace(d)][faceQ(P, loFace(d), i, j, k)] += (double)g;
            }
            if (idx == n - 1) {
                const float ub = uf[faceLoAt(P, d, i, j, k, 1)];
                const float g = ub < 0 ? rA * (A.abar[c + N * hi] - A.abar[c]) : 0.0f;
                if (P.bct[hiFace(d)] == IFL_OUTLET) o += g;
                else if (P.bct[hiFace(d)] == IFL_INLET) A.gprof[hiFace(d)][faceQ(P, hiFace(d), i, j, k)] += (double)g;
            }
            const float nu = o + wOf(A.gu[d], c);
            const double e = (double)nu - (double)A.lu[d][c];
            red[d] += e * e;
            A.lu[d][c] = nu;
            if (idx >= 1) {
                const int fi = faceLoAt(P, d, i, j, k, 0);
                const int Lc = c - st;
                const float ufv = uf[fi];
                float a = 0.0f;
                if (ufv < 0) a += A.abar[Lc + N * hi] - A.abar[Lc];
                if (ufv > 0) a += A.abar[c] - A.abar[c + N * lo];
                const float nf = P.rho * areaOf(P, d) * a;
                const double ef = (double)nf - (double)A.lf[d][fi];
                red[3 + d] += ef * ef;
                A.lf[d][fi] = nf;
            }
        }
    });
    storeBlockSums<6>(red, A.part + 2 * A.nblk + 6 * flatBlock());
}

// ---------------- monitor: increments, scaled increment, mu-bar ----------------
__global__ __launch_bounds__(kBlock) void k_adj_reduce(AdjBufs A) {
    double a[2], b[6];
    sumPartials<2>(A.part, A.nblk, a);
    sumPartials<6>(A.part + 2 * A.nblk, A.nblk, b);
    if (threadIdx.x == 0) {
        AdjState &S = *A.st;
        const double inc[7] = {b[0], b[1], b[2], a[1], b[3], b[4], b[5]};
        double sc = 0.0;
        for (int q = 0; q < 7; ++q) {
            S.inc[q] = inc[q];
            if (S.iters < 5) S.incMax[q] = fmax(S.incMax[q], inc[q]);
            if (S.incMax[q] > 0) sc = fmax(sc, sqrt(inc[q] / S.incMax[q]));
        }
        S.scaled = sc;
        S.mub = a[0];
        S.iters += 1;
    }
}

// ---------------- launchers ----------------
static dim3 adjGrid(const CaseParams &P) {
    return dim3((unsigned)((P.nx + 63) / 64), (unsigned)((P.ny * P.nz + 4 * kRows - 1) / (4 * kRows)));
}

#define IFL_ADJ_LAUNCH(L, KERNEL, ...)                                                                   \
    do {                                                                                                \
        const CaseParams &P_ = *(L).P;                                                                  \
        if (P_.dim == 3) KERNEL<3><<<adjGrid(P_), kBlock, 0, (L).s>>>(P_, ##__VA_ARGS__);                \
        else KERNEL<2><<<adjGrid(P_), kBlock, 0, (L).s>>>(P_, ##__VA_ARGS__);                            \
    } while (0)

void launch_adj_pcorr(const Lc &L, const AdjBufs &A) { IFL_ADJ_LAUNCH(L, k_adj_pcorr, A); }
void launch_adj_sweep_t(const Lc &L, const AdjBufs &A, const float *in, float *out, float *bb, const float *coef,
                        float rlx, int col) {
    (void)A;
    IFL_ADJ_LAUNCH(L, k_adj_sweep_t, in, out, bb, coef, rlx, col);
}
void launch_adj_rc(const Lc &L, const AdjBufs &A) { IFL_ADJ_LAUNCH(L, k_adj_rc, A); }
void launch_adj_coef(const Lc &L, const AdjBufs &A) { IFL_ADJ_LAUNCH(L, k_adj_coef, A); }
void launch_adj_face(const Lc &L, const AdjBufs &A) { IFL_ADJ_LAUNCH(L, k_adj_face, A); }
void launch_adj_reduce(const Lc &L, const AdjBufs &A) { k_adj_reduce<<<1, kBlock, 0, L.s>>>(A); }

}  // namespace ifl
