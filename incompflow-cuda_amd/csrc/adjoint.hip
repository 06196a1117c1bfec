// adjoint.hip - discrete adjoint of the converged SIMPLE state on gfx950.
//
// One thread per cell (the cell grid of kernels.hip).  Each kernel computes, for its cell, the
// adjoint elements that land on that cell as a sum over the cells and faces that reach it, in
// the fixed order of oracle/adjoint_oracle.hpp, so no atomics are needed and the result is
// deterministic.  The reference has no adjoint; the stages transposed are those of one outer
// iteration of Solver.cu:160-290 (kernels.cuh:97-843) at its fixed point.
#include "ifl_adjoint.hpp"
#include "ifl_stencil.hpp"

#include <utility>

namespace ifl {

namespace {

// rows per thread as the forward cell kernels (P.cellRows: 1 on grids up to 2^19 cells, where a
// launch is latency bound, kRows above)
template <class F>
__device__ __forceinline__ void forCellsA(const CaseParams &P, F &&f) {
    const int rows = P.ny * P.nz;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int R = P.cellRows;
    const int r0 = blockIdx.y * (4 * R) + (threadIdx.x >> 6);
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

// ---------------- ns (<= K) transposed red-black sweeps per launch, 2D ----------------
// The transposed half sweep is local: a cell of the half sweep's colour hands its S to the source
// adjoint and keeps phb - rlx phb; every other cell subtracts, per neighbour, the neighbour's
// coefficient towards it times the neighbour's S (k_adj_sweep_t).  A workgroup owns a TX x TY
// tile and keeps the tile plus a 2K-cell halo in registers -- per cell phb, aC, the coefficients
// of its four neighbours towards it and the source adjoint, for the two cells (one of each
// colour) of each of its pairs -- and exchanges only S through LDS, one barrier per half sweep.
// Half sweep h leaves the outer h rings of the staged region stale (their neighbours are not
// staged), so after 2K half sweeps the tile is exact: each cell goes through the arithmetic of
// k_adj_sweep_t, in the same order, and the result is bit-identical.
template <int TX, int TY, int K>
__global__ __launch_bounds__(kBlock) void k_adj_tile2d(CaseParams P, const float *in, float *out, float *bb,
                                                       const float *coef, float rlx, int ns) {
    constexpr int H = 2 * K;
    constexpr int HX = TX + 2 * H, HY = TY + 2 * H, HX2 = HX / 2;
    constexpr int NPAIR = HX2 * HY;
    constexpr int NP = (NPAIR + kBlock - 1) / kBlock;   // pairs per thread
    constexpr int WS = HX + 2;                          // LDS row: the staged row and a pad cell each side
    constexpr int DUMMY = (HY + 2) * WS;                // S slot of the pairs past the region
    static_assert(TX % 2 == 0 && TY % 2 == 0, "even tile origins keep the global red-black colours");
    __shared__ float lS[DUMMY + 1];
    const int nx = P.nx, ny = P.ny, N = P.N;
    const int tilesX = (nx + TX - 1) / TX;
    const int by = (int)blockIdx.x / tilesX, bx = (int)blockIdx.x - by * tilesX;
    const int ox = bx * TX - H, oy = by * TY - H;       // origin of the staged region
    const int t = threadIdx.x;

    // [m][c]: the colour-c cell of the thread's m-th pair
    float ph[NP][2], ac[NP][2], kE[NP][2], kW[NP][2], kN[NP][2], kS[NP][2], sb[NP][2];
    int li[NP][2];          // LDS slot of the cell's S
    unsigned fl[NP][2];     // neighbours in the grid: 1 east, 2 west, 4 north, 8 south; 16 owned cell
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        const int e = t + kBlock * m;
        const int ly = e / HX2, px = e - ly * HX2;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            ph[m][c] = 0.0f;
            ac[m][c] = 1.0f;
            kE[m][c] = kW[m][c] = kN[m][c] = kS[m][c] = 0.0f;
            sb[m][c] = 0.0f;
            fl[m][c] = 0u;
            li[m][c] = DUMMY;
            if (NPAIR % kBlock == 0 || e < NPAIR) {
                const int lx = 2 * px + (c ^ (ly & 1));   // colour (lx + ly) & 1 == c
                li[m][c] = (ly + 1) * WS + lx + 1;
                const int gx = ox + lx, gy = oy + ly;
                if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) {
                    const int g = gx + nx * gy;
                    ph[m][c] = in[g];
                    ac[m][c] = coef[g];
                    unsigned f = 0u;
                    if (gx < nx - 1) { f |= 1u; kE[m][c] = coef[g + 1 + N * sW]; }
                    if (gx > 0) { f |= 2u; kW[m][c] = coef[g - 1 + N * sE]; }
                    if (gy < ny - 1) { f |= 4u; kN[m][c] = coef[g + nx + N * sS]; }
                    if (gy > 0) { f |= 8u; kS[m][c] = coef[g - nx + N * sN]; }
                    if (lx >= H && lx < H + TX && ly >= H && ly < H + TY) {
                        f |= 16u;
                        sb[m][c] = bb[g];
                    }
                    fl[m][c] = f;
                }
            }
        }
    }
    // the pad ring is read only by stale cells; zero it so they stay finite
    for (int r = t; r < 2 * WS + 2 * HY; r += kBlock) {
        int s;
        if (r < WS) s = r;
        else if (r < 2 * WS) s = (HY + 1) * WS + (r - WS);
        else if (r < 2 * WS + HY) s = (r - 2 * WS + 1) * WS;
        else s = (r - 2 * WS - HY + 1) * WS + WS - 1;
        lS[s] = 0.0f;
    }

    auto half = [&](auto colc) {
        constexpr int col = decltype(colc)::value, oc = col ^ 1;
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const float x = ph[m][col];
            const float S = rlx * x / ac[m][col];
            lS[li[m][col]] = S;
            sb[m][col] += S;
            ph[m][col] = x - rlx * x;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NP; ++m) {
            const int l = li[m][oc];
            const unsigned f = fl[m][oc];
            float acc = ph[m][oc];
            if (f & 1u) acc -= kE[m][oc] * lS[l + 1];
            if (f & 2u) acc -= kW[m][oc] * lS[l - 1];
            if (f & 4u) acc -= kN[m][oc] * lS[l + WS];
            if (f & 8u) acc -= kS[m][oc] * lS[l - WS];
            ph[m][oc] = acc;
        }
    };
    for (int s = 0; s < ns; ++s) {
        half(std::integral_constant<int, 1>{});
        half(std::integral_constant<int, 0>{});
    }
#pragma unroll
    for (int m = 0; m < NP; ++m)
#pragma unroll
        for (int c = 0; c < 2; ++c)
            if (fl[m][c] & 16u) {
                const int ly = li[m][c] / WS - 1, lx = li[m][c] - (ly + 1) * WS - 1;
                const int g = (ox + lx) + nx * (oy + ly);
                out[g] = ph[m][c];
                bb[g] = sb[m][c];
            }
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
    const int R = P.cellRows;
    return dim3((unsigned)((P.nx + 63) / 64), (unsigned)((P.ny * P.nz + 4 * R - 1) / (4 * R)));
}

int adj_blocks(const CaseParams &P) {
    const dim3 g = adjGrid(P);
    return (int)(g.x * g.y);
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
// 2D transposed sweeps: tile shapes and sweeps per launch
namespace {
template <int TX, int TY, int K>
void adjTile(const Lc &L, const float *in, float *out, float *bb, const float *coef, float rlx, int ns) {
    const CaseParams &P = *L.P;
    const unsigned tiles = (unsigned)(((P.nx + TX - 1) / TX) * ((P.ny + TY - 1) / TY));
    k_adj_tile2d<TX, TY, K><<<tiles, kBlock, 0, L.s>>>(P, in, out, bb, coef, rlx, ns);
}
}  // namespace

int adj_tile_sweeps(int shape) {
    static constexpr int k[kAdjShapes] = {4, 4, 2, 4, 2, 4};
    return shape >= 0 && shape < kAdjShapes ? k[shape] : 0;
}

void launch_adj_tile2d(const Lc &L, int shape, const float *in, float *out, float *bb, const float *coef, float rlx,
                       int ns) {
    switch (shape) {
    case 0: adjTile<64, 16, 4>(L, in, out, bb, coef, rlx, ns); break;
    case 1: adjTile<32, 16, 4>(L, in, out, bb, coef, rlx, ns); break;
    case 2: adjTile<64, 8, 2>(L, in, out, bb, coef, rlx, ns); break;
    case 3: adjTile<32, 8, 4>(L, in, out, bb, coef, rlx, ns); break;
    case 4: adjTile<64, 16, 2>(L, in, out, bb, coef, rlx, ns); break;
    default: adjTile<64, 8, 4>(L, in, out, bb, coef, rlx, ns); break;
    }
}

void launch_adj_rc(const Lc &L, const AdjBufs &A) { IFL_ADJ_LAUNCH(L, k_adj_rc, A); }
void launch_adj_coef(const Lc &L, const AdjBufs &A) { IFL_ADJ_LAUNCH(L, k_adj_coef, A); }
void launch_adj_face(const Lc &L, const AdjBufs &A) { IFL_ADJ_LAUNCH(L, k_adj_face, A); }
void launch_adj_reduce(const Lc &L, const AdjBufs &A) { k_adj_reduce<<<1, kBlock, 0, L.s>>>(A); }

}  // namespace ifl
