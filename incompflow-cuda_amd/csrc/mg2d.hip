// mg2d.hip - gfx950 kernels of the multigrid p' solve on 2D grids (IFL_PSOLVE_MG).
//
// Per outer iteration: k_mg_coarsen2d builds the coarse matrices down to the bottom kernel's first
// level (one pass makes up to six levels from 64x64 tiles of its source level, each tile summing
// its own cells hierarchically in LDS; the couplings of a tile's west and south edge cells towards
// -x and -y come from the same sums over the neighbouring column and row).  Per V-cycle: k_mg_down2d on every level above the bottom
// (NU red-black sweeps from x = 0 on a 64x32 or 32x16 tile staged with a (2NU+1)-cell halo, then the FP64
// residual of the tile and its restriction into the coarse b), k_mg_bottom2d for the bottom levels
// in one workgroup, and k_mg_up2d on the way back (x += over * x_parent on the staged tile with a
// 2NU halo, NU sweeps).  Halo cells are recomputed from the same values their owning tile uses, so
// every tile's cells are bit-identical to a level-wide sweep; the arithmetic is the oracle's
// (simple_oracle.hpp, mgBuild .. mgCycle) expression by expression (-ffp-contract=off).
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_mg.hpp"

#include <algorithm>

namespace ifl {

// Level-0 outlet term of cell (i, j): the coupling rho A^2 / aC to the ghost p' = 0 beyond every
// outlet face the cell touches, faces in the order E, W, N, S (1/aC is P.rinv, k_pcorr_assemble).
__device__ __forceinline__ float mgOutlet0(const CaseParams &P, int i, int j, int g) {
    const bool atE = i == P.nx - 1 && P.bct[E] == IFL_OUTLET, atW = i == 0 && P.bct[W] == IFL_OUTLET;
    const bool atN = j == P.ny - 1 && P.bct[N] == IFL_OUTLET, atS = j == 0 && P.bct[S] == IFL_OUTLET;
    float dg = 0.0f;
    if (atE | atW | atN | atS) {
        const float rinv = P.rinv[g];
        if (atE) dg += P.rho * P.Ax * P.Ax * rinv;
        if (atW) dg += P.rho * P.Ax * P.Ax * rinv;
        if (atN) dg += P.rho * P.Ay * P.Ay * rinv;
        if (atS) dg += P.rho * P.Ay * P.Ay * rinv;
    }
    return dg;
}

// ---------------- coarse matrices ----------------
// One workgroup per 64x64 tile of the source level; levels src+1 .. src+m of the tile's footprint.
template <bool L0>
__global__ __launch_bounds__(256) void k_mg_coarsen2d(CaseParams P, MgCoarsen G) {
    constexpr int T = 64, NT = 256;
    __shared__ float cE[2][T * T], cN[2][T * T], cD[2][T * T];
    __shared__ float hW[2][T], hS[2][T];     // couplings of the west column's +x and south row's +y
    if (caseDone(P)) return;
    const int tid = threadIdx.x;
    const int nxs = L0 ? P.nx : G.s.nx, nys = L0 ? P.ny : G.s.ny;
    const int tilesX = (nxs + T - 1) / T;
    const int t = xcdSwizzle((int)blockIdx.x, (int)gridDim.x);
    const int by = t / tilesX, bx = t - by * tilesX;
    const int x0 = bx * T, y0 = by * T;
    const float *srcE = L0 ? P.pcoef + (size_t)P.N * sE : G.s.E;
    const float *srcN = L0 ? P.pcoef + (size_t)P.N * sN : G.s.N;
    for (int i = tid; i < T * T; i += NT) {
        const int ly = i / T, lx = i - ly * T, gx = x0 + lx, gy = y0 + ly;
        float e = 0.0f, n = 0.0f, d = 0.0f;
        if (gx < nxs && gy < nys) {
            const int g = gy * nxs + gx;
            e = srcE[g];
            n = srcN[g];
            d = L0 ? mgOutlet0(P, gx, gy, g) : G.s.dg[g];
        }
        cE[0][i] = e;
        cN[0][i] = n;
        cD[0][i] = d;
    }
    if (tid < T) {
        float hw = 0.0f, hs = 0.0f;
        if (x0 > 0 && y0 + tid < nys) hw = srcE[(y0 + tid) * nxs + x0 - 1];
        if (y0 > 0 && x0 + tid < nxs) hs = srcN[(y0 - 1) * nxs + x0 + tid];
        hW[0][tid] = hw;
        hS[0][tid] = hs;
    }
    __syncthreads();
    int nxp = nxs, nyp = nys, w = T, px0 = x0, py0 = y0;
#pragma unroll
    for (int q = 0; q < kMgCoarsenDepth; ++q) {
        if (q < G.m) {
            const int cur = q & 1, nxt = cur ^ 1;
            const int wc = w / 2, X0 = px0 / 2, Y0 = py0 / 2;
            for (int i = tid; i < wc * wc; i += NT) {
                const int J = i / wc, I = i - J * wc;
                const bool hasI1 = 2 * (X0 + I) + 1 < nxp, hasJ1 = 2 * (Y0 + J) + 1 < nyp;
                const int i0 = 2 * I, i1 = hasI1 ? i0 + 1 : i0, j0 = 2 * J, j1 = hasJ1 ? j0 + 1 : j0;
                float e = 0.0f, n = 0.0f, d = 0.0f;
                e += cE[cur][j0 * w + i1];
                if (hasJ1) e += cE[cur][j1 * w + i1];
                n += cN[cur][j1 * w + i0];
                if (hasI1) n += cN[cur][j1 * w + i1];
                d += cD[cur][j0 * w + i0];
                if (hasI1) d += cD[cur][j0 * w + i1];
                if (hasJ1) {
                    d += cD[cur][j1 * w + i0];
                    if (hasI1) d += cD[cur][j1 * w + i1];
                }
                cE[nxt][i] = e;
                cN[nxt][i] = n;
                cD[nxt][i] = d;
            }
            if (tid < wc) {
                const bool hasJ1 = 2 * (Y0 + tid) + 1 < nyp, hasI1 = 2 * (X0 + tid) + 1 < nxp;
                float hw = 0.0f, hs = 0.0f;
                hw += hW[cur][2 * tid];
                if (hasJ1) hw += hW[cur][2 * tid + 1];
                hs += hS[cur][2 * tid];
                if (hasI1) hs += hS[cur][2 * tid + 1];
                hW[nxt][tid] = hw;
                hS[nxt][tid] = hs;
            }
            __syncthreads();
            const MgLevel &D = G.d[q];
            for (int i = tid; i < wc * wc; i += NT) {
                const int J = i / wc, I = i - J * wc, gI = X0 + I, gJ = Y0 + J;
                if (gI < D.nx && gJ < D.ny) {
                    const float e = cE[nxt][i], n = cN[nxt][i], d = cD[nxt][i];
                    const float ew = I > 0 ? cE[nxt][i - 1] : (gI > 0 ? hW[nxt][J] : 0.0f);
                    const float ns = J > 0 ? cN[nxt][i - wc] : (gJ > 0 ? hS[nxt][I] : 0.0f);
                    const float cc = -(e + ew + n + ns);
                    const int g = gJ * D.nx + gI;
                    D.C[g] = cc + d;
                    D.E[g] = e;
                    D.N[g] = n;
                    D.dg[g] = d;
                }
            }
            nxp = D.nx;
            nyp = D.ny;
            w = wc;
            px0 = X0;
            py0 = Y0;
        }
    }
}

// ---------------- tiles of the levels above the bottom ----------------
// Staged layout: colour split, cell (lx, ly) of the (TX + 2H) x (TY + 2H) region at
// [(lx + ly) & 1][ly][lx / 2]; tile origins are even, so (lx + ly) & 1 is the global colour.
template <int TX, int TY, int H>
struct MgTile {
    static constexpr int HX = TX + 2 * H, HY = TY + 2 * H, HX2 = HX / 2, CO = HY * HX2;
    static_assert(HX % 2 == 0, "colour split needs an even width");
    __device__ static int at(int lx, int ly) { return ((lx + ly) & 1) * CO + ly * HX2 + (lx >> 1); }
};

// Load the region around tile (x0, y0) of a level: x (zero, or the level's x plus over times the
// parent's), b, centre, +x and +y couplings; off-grid cells are 0.
template <int TX, int TY, int H, int NT, bool L0, bool XZ, bool PROLONG>
__device__ __forceinline__ void mgStage(const CaseParams &P, const MgLevel &F, const MgLevel &Cc, const float *xg,
                                        float over, int x0, int y0, int nx, int ny, float *tx, float *tb, float *tc,
                                        float *te, float *tn) {
    using Tl = MgTile<TX, TY, H>;
    const float *bg = L0 ? P.bp : F.b;
    for (int i = threadIdx.x; i < Tl::HX * Tl::HY; i += NT) {
        const int ly = i / Tl::HX, lx = i - ly * Tl::HX;
        const int gx = x0 - H + lx, gy = y0 - H + ly;
        float xv = 0.0f, bv = 0.0f, cv = 0.0f, ev = 0.0f, nv = 0.0f;
        if ((unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny) {
            const int g = gy * nx + gx;
            if (!XZ) xv = xg[g];
            if (PROLONG) xv = xv + over * Cc.xu[(gy >> 1) * Cc.nx + (gx >> 1)];
            bv = bg[g];
            if (L0) {
                ev = P.pcoef[g + (size_t)P.N * ifl::sE];
                nv = P.pcoef[g + (size_t)P.N * ifl::sN];
                cv = P.pcoef[g] + mgOutlet0(P, gx, gy, g);
            } else {
                ev = F.E[g];
                nv = F.N[g];
                cv = F.C[g];
            }
        }
        const int a = Tl::at(lx, ly);
        tx[a] = xv;
        tb[a] = bv;
        tc[a] = cv;
        te[a] = ev;
        tn[a] = nv;
    }
}

// Half sweep of colour col over the cells within R of the tile (staged coordinates).
template <int TX, int TY, int H, int NT>
__device__ __forceinline__ void mgHalf(float *tx, const float *tb, const float *tc, const float *te, const float *tn,
                                       int R, int col, int x0, int y0, int nx, int ny) {
    using Tl = MgTile<TX, TY, H>;
    const int l0 = H - R, W2 = (TX + 2 * R) / 2, cnt = W2 * (TY + 2 * R);
    for (int i = threadIdx.x; i < cnt; i += NT) {
        const int row = i / W2, k = i - row * W2;
        const int ly = l0 + row;
        const int lx = l0 + 2 * k + ((col + ly + l0) & 1);
        const int gx = x0 - H + lx, gy = y0 - H + ly;
        if ((unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny) {
            const int o = col * Tl::CO + ly * Tl::HX2 + (lx >> 1);
            const int ot = (1 - col) * Tl::CO + ly * Tl::HX2 + (lx >> 1);
            const int w = ot + (lx & 1) - 1, e = w + 1, n = ot + Tl::HX2, s = ot - Tl::HX2;
            const float v = tb[o] - te[o] * tx[e] - te[w] * tx[w] - tn[o] * tx[n] - tn[s] * tx[s];
            tx[o] = v / tc[o];
        }
    }
}

// Down: NU sweeps (x from zero on the first cycle, or from p' on level 0 of a later cycle), the
// tile's FP64 residual, its x, and the restriction into the coarse b.
template <int TX, int TY, int NT, int NU, bool L0, bool XZ>
__global__ __launch_bounds__(NT) void k_mg_down2d(CaseParams P, MgLevel F, MgLevel Cc) {
    constexpr int H = 2 * NU + 1;
    using Tl = MgTile<TX, TY, H>;
    __shared__ float tx[2 * Tl::CO], tb[2 * Tl::CO], tc[2 * Tl::CO], te[2 * Tl::CO], tn[2 * Tl::CO];
    __shared__ double sr[TX * TY];
    if (caseDone(P)) return;
    const int nx = L0 ? P.nx : F.nx, ny = L0 ? P.ny : F.ny;
    const int tilesX = (nx + TX - 1) / TX;
    const int t = xcdSwizzle((int)blockIdx.x, (int)gridDim.x);
    const int by = t / tilesX, bx = t - by * tilesX;
    const int x0 = bx * TX, y0 = by * TY;
    // reads p' (a later cycle) or nothing; writes the down-pass x (another array than it reads)
    mgStage<TX, TY, H, NT, L0, XZ, false>(P, F, Cc, L0 ? P.pc : F.x, 0.0f, x0, y0, nx, ny, tx, tb, tc, te, tn);
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 2 * NU; ++p) {
        mgHalf<TX, TY, H, NT>(tx, tb, tc, te, tn, H - 1 - p, p & 1, x0, y0, nx, ny);
        __syncthreads();
    }
    float *xg = L0 ? P.alt[3] : F.x;
    for (int i = threadIdx.x; i < TX * TY; i += NT) {
        const int r = i / TX, xx = i - r * TX, gx = x0 + xx, gy = y0 + r;
        double res = 0.0;
        if (gx < nx && gy < ny) {
            const int lx = xx + H, ly = r + H, col = (lx + ly) & 1;
            const int o = col * Tl::CO + ly * Tl::HX2 + (lx >> 1);
            const int ot = (1 - col) * Tl::CO + ly * Tl::HX2 + (lx >> 1);
            const int w = ot + (lx & 1) - 1, e = w + 1, n = ot + Tl::HX2, s = ot - Tl::HX2;
            const float xo = tx[o];
            double q = (double)tb[o];
            q -= (double)tc[o] * (double)xo;
            q -= (double)te[o] * (double)tx[e];
            q -= (double)te[w] * (double)tx[w];
            q -= (double)tn[o] * (double)tx[n];
            q -= (double)tn[s] * (double)tx[s];
            res = q;
            xg[gy * nx + gx] = xo;
        }
        sr[i] = res;
    }
    __syncthreads();
    constexpr int CX = TX / 2, CY = TY / 2;
    for (int i = threadIdx.x; i < CX * CY; i += NT) {
        const int J = i / CX, I = i - J * CX, gI = x0 / 2 + I, gJ = y0 / 2 + J;
        if (gI < Cc.nx && gJ < Cc.ny) {
            const bool hasI1 = 2 * gI + 1 < nx, hasJ1 = 2 * gJ + 1 < ny;
            double acc = 0.0;
            acc += sr[(2 * J) * TX + 2 * I];
            if (hasI1) acc += sr[(2 * J) * TX + 2 * I + 1];
            if (hasJ1) {
                acc += sr[(2 * J + 1) * TX + 2 * I];
                if (hasI1) acc += sr[(2 * J + 1) * TX + 2 * I + 1];
            }
            Cc.b[gJ * Cc.nx + gI] = (float)acc;
        }
    }
}

// Up: x += over * x_parent on the staged region, NU sweeps, the tile's x.
template <int TX, int TY, int NT, int NU, bool L0>
__global__ __launch_bounds__(NT) void k_mg_up2d(CaseParams P, MgLevel F, MgLevel Cc, float over) {
    constexpr int H = 2 * NU;
    using Tl = MgTile<TX, TY, H>;
    __shared__ float tx[2 * Tl::CO], tb[2 * Tl::CO], tc[2 * Tl::CO], te[2 * Tl::CO], tn[2 * Tl::CO];
    if (caseDone(P)) return;
    const int nx = L0 ? P.nx : F.nx, ny = L0 ? P.ny : F.ny;
    const int tilesX = (nx + TX - 1) / TX;
    const int t = xcdSwizzle((int)blockIdx.x, (int)gridDim.x);
    const int by = t / tilesX, bx = t - by * tilesX;
    const int x0 = bx * TX, y0 = by * TY;
    // reads the down-pass x around the tile and the parents' up-pass x; writes the up-pass x
    mgStage<TX, TY, H, NT, L0, false, true>(P, F, Cc, L0 ? P.alt[3] : F.x, over, x0, y0, nx, ny, tx, tb, tc, te, tn);
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 2 * NU; ++p) {
        mgHalf<TX, TY, H, NT>(tx, tb, tc, te, tn, H - 1 - p, p & 1, x0, y0, nx, ny);
        __syncthreads();
    }
    float *xg = L0 ? P.pc : F.xu;
    for (int i = threadIdx.x; i < TX * TY; i += NT) {
        const int r = i / TX, xx = i - r * TX, gx = x0 + xx, gy = y0 + r;
        if (gx < nx && gy < ny) xg[gy * nx + gx] = tx[Tl::at(xx + H, r + H)];
    }
}

// ---------------- the bottom levels, one workgroup ----------------
// Levels lb .. nl-1 in LDS, six row-major planes each (x, b, C, E, N, outlet term).  The kernel
// coarsens the matrices of levels lb+1.. itself (the whole level is in LDS, so the couplings of a
// cell towards -x and -y are its neighbours' +x and +y ones directly), then runs the rest of the
// V-cycle.  Threads form 32 columns x NT/32 rows; thread (tx, ty) takes the cells of colour col in
// rows ty, ty + NT/32, ... at i = ((col + j) & 1) + 2 tx + 64 m: no division by the level width.
constexpr int kMgBottomFloats = 6 * 5600;
constexpr int kMgPlanes = 6;

// `sweeps` red-black sweeps of one level in LDS.  With Lc (the coarser level's x, cnx wide), the
// first (red) half sweep reads its neighbours as x + over * x_parent: that is the coarse correction
// of the black cells, which nothing else reads before the black half sweep overwrites them, and the
// red cells' own corrected values are never read (an unrelaxed update does not use the old value),
// so this equals correcting every cell first and sweeping after, bit for bit, without its pass.
template <int NT>
__device__ __forceinline__ void mgSmoothLds(float *L, int nx, int ny, int sweeps, const float *Lc = nullptr,
                                            int cnx = 0, float over = 0.0f) {
    constexpr int TY = NT / 32;
    const int n = nx * ny, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    float *x = L;
    const float *b = L + n, *C = L + 2 * n, *E = L + 3 * n, *Nn = L + 4 * n;
    for (int s = 0; s < sweeps; ++s)
        for (int col = 0; col < 2; ++col) {
            const bool pro = Lc != nullptr && s == 0 && col == 0;
            for (int j = ty; j < ny; j += TY)
                for (int i = ((col + j) & 1) + 2 * tx; i < nx; i += 64) {
                    const int c = i + nx * j;
                    float xE = i < nx - 1 ? x[c + 1] : 0.0f, xW = i > 0 ? x[c - 1] : 0.0f;
                    float xN = j < ny - 1 ? x[c + nx] : 0.0f, xS = j > 0 ? x[c - nx] : 0.0f;
                    if (pro) {
                        if (i < nx - 1) xE = xE + over * Lc[(j >> 1) * cnx + ((i + 1) >> 1)];
                        if (i > 0) xW = xW + over * Lc[(j >> 1) * cnx + ((i - 1) >> 1)];
                        if (j < ny - 1) xN = xN + over * Lc[((j + 1) >> 1) * cnx + (i >> 1)];
                        if (j > 0) xS = xS + over * Lc[((j - 1) >> 1) * cnx + (i >> 1)];
                    }
                    const float aW = i > 0 ? E[c - 1] : 0.0f, aS = j > 0 ? Nn[c - nx] : 0.0f;
                    const float v = b[c] - E[c] * xE - aW * xW - Nn[c] * xN - aS * xS;
                    x[c] = v / C[c];
                }
            __syncthreads();
        }
}

__device__ __forceinline__ double mgResidualLds(const float *L, int nx, int ny, int i, int j) {
    const int n = nx * ny, c = i + nx * j;
    const float *x = L, *b = L + n, *C = L + 2 * n, *E = L + 3 * n, *Nn = L + 4 * n;
    double r = (double)b[c];
    r -= (double)C[c] * (double)x[c];
    r -= (double)E[c] * (double)(i < nx - 1 ? x[c + 1] : 0.0f);
    r -= (double)(i > 0 ? E[c - 1] : 0.0f) * (double)(i > 0 ? x[c - 1] : 0.0f);
    r -= (double)Nn[c] * (double)(j < ny - 1 ? x[c + nx] : 0.0f);
    r -= (double)(j > 0 ? Nn[c - nx] : 0.0f) * (double)(j > 0 ? x[c - nx] : 0.0f);
    return r;
}

// recordSweeps > 0 (the first cycle): the p' solve's sweep count (fine-level sweeps) for the record.
template <int NU, bool L0, bool XZ, int NT>
__global__ __launch_bounds__(NT) void k_mg_bottom2d(CaseParams P, const MgLevel *lv, int lb, int nl, int ncoarse,
                                                    float over, int recordSweeps) {
    constexpr int TY = NT / 32;
    __shared__ float lds[kMgBottomFloats];
    if (caseDone(P)) return;
    const int tid = threadIdx.x, tx = tid & 31, ty = tid >> 5;
    if (recordSweeps > 0 && tid == 0) {
        P.st->sm[3].sweeps = recordSweeps;
        P.st->sm[3].lastSum = 0.0;
        P.st->sweepsIter[3] += recordSweeps;
    }
    // level lb from global memory (level 0: from the case's p' system)
    {
        const int nx = lv[lb].nx, n = nx * lv[lb].ny;
        float *L = lds;
        for (int c = tid; c < n; c += NT) {
            float xv = 0.0f, bv, cv, ev, nv, dv;
            if (L0) {
                if (!XZ) xv = P.pc[c];
                bv = P.bp[c];
                ev = P.pcoef[c + (size_t)P.N * sE];
                nv = P.pcoef[c + (size_t)P.N * sN];
                dv = mgOutlet0(P, c % nx, c / nx, c);
                cv = P.pcoef[c] + dv;
            } else {
                if (!XZ) xv = lv[lb].x[c];
                bv = lv[lb].b[c];
                cv = lv[lb].C[c];
                ev = lv[lb].E[c];
                nv = lv[lb].N[c];
                dv = lv[lb].dg[c];
            }
            L[c] = xv;
            L[n + c] = bv;
            L[2 * n + c] = cv;
            L[3 * n + c] = ev;
            L[4 * n + c] = nv;
            L[5 * n + c] = dv;
        }
    }
    __syncthreads();
    // the coarser levels' matrices (mgBuild's sums, in its order), x = 0
    int base = 0;
    for (int l = lb; l < nl - 1; ++l) {
        const int nx = lv[l].nx, ny = lv[l].ny, n = nx * ny;
        const int cnx = lv[l + 1].nx, cny = lv[l + 1].ny, cn = cnx * cny;
        const float *F = lds + base;
        float *Lc = lds + base + kMgPlanes * n;
        for (int J = ty; J < cny; J += TY)
            for (int I = tx; I < cnx; I += 32) {
                const bool hasI1 = 2 * I + 1 < nx, hasJ1 = 2 * J + 1 < ny;
                const int i0 = 2 * I, i1 = hasI1 ? i0 + 1 : i0, j0 = 2 * J, j1 = hasJ1 ? j0 + 1 : j0;
                const float *fE = F + 3 * n, *fN = F + 4 * n, *fD = F + 5 * n;
                float e = 0.0f, nn = 0.0f, d = 0.0f;
                e += fE[i1 + nx * j0];
                if (hasJ1) e += fE[i1 + nx * j1];
                nn += fN[i0 + nx * j1];
                if (hasI1) nn += fN[i1 + nx * j1];
                d += fD[i0 + nx * j0];
                if (hasI1) d += fD[i1 + nx * j0];
                if (hasJ1) {
                    d += fD[i0 + nx * j1];
                    if (hasI1) d += fD[i1 + nx * j1];
                }
                const int C = I + cnx * J;
                Lc[C] = 0.0f;
                Lc[3 * cn + C] = e;
                Lc[4 * cn + C] = nn;
                Lc[5 * cn + C] = d;
            }
        __syncthreads();
        for (int J = ty; J < cny; J += TY)
            for (int I = tx; I < cnx; I += 32) {
                const int C = I + cnx * J;
                const float e = Lc[3 * cn + C], nn = Lc[4 * cn + C];
                const float ew = I > 0 ? Lc[3 * cn + C - 1] : 0.0f, ns = J > 0 ? Lc[4 * cn + C - cnx] : 0.0f;
                const float cc = -(e + ew + nn + ns);
                Lc[2 * cn + C] = cc + Lc[5 * cn + C];
            }
        __syncthreads();
        base += kMgPlanes * n;
    }
    // down
    base = 0;
    for (int l = lb; l < nl - 1; ++l) {
        const int nx = lv[l].nx, ny = lv[l].ny, n = nx * ny;
        const int cnx = lv[l + 1].nx, cny = lv[l + 1].ny, cn = cnx * cny;
        float *L = lds + base, *Lc = lds + base + kMgPlanes * n;
        mgSmoothLds<NT>(L, nx, ny, NU);
        for (int J = ty; J < cny; J += TY)
            for (int I = tx; I < cnx; I += 32) {
                const bool hasI1 = 2 * I + 1 < nx, hasJ1 = 2 * J + 1 < ny;
                double acc = 0.0;
                acc += mgResidualLds(L, nx, ny, 2 * I, 2 * J);
                if (hasI1) acc += mgResidualLds(L, nx, ny, 2 * I + 1, 2 * J);
                if (hasJ1) {
                    acc += mgResidualLds(L, nx, ny, 2 * I, 2 * J + 1);
                    if (hasI1) acc += mgResidualLds(L, nx, ny, 2 * I + 1, 2 * J + 1);
                }
                Lc[cn + I + cnx * J] = (float)acc;
            }
        __syncthreads();
        base += kMgPlanes * n;
    }
    mgSmoothLds<NT>(lds + base, lv[nl - 1].nx, lv[nl - 1].ny, ncoarse);
    // up: the coarse correction folded into the first half sweep (mgSmoothLds)
    for (int l = nl - 2; l >= lb; --l) {
        const int nx = lv[l].nx, ny = lv[l].ny, n = nx * ny, cnx = lv[l + 1].nx;
        const float *Lc = lds + base;
        base -= kMgPlanes * n;
        mgSmoothLds<NT>(lds + base, nx, ny, NU, Lc, cnx, over);
    }
    {
        const int n = lv[lb].nx * lv[lb].ny;
        float *xg = L0 ? P.pc : lv[lb].xu;      // what the up pass reads as parents
        for (int c = tid; c < n; c += NT) xg[c] = lds[c];
    }
}

// ---------------- host side ----------------
int mg_levels(int nx, int ny, int (*dims)[2]) {
    int nl = 0;
    dims[nl][0] = nx;
    dims[nl][1] = ny;
    ++nl;
    while (nl < kMgMaxLevels && dims[nl - 1][0] >= 4 && dims[nl - 1][1] >= 4) {
        dims[nl][0] = (dims[nl - 1][0] + 1) / 2;
        dims[nl][1] = (dims[nl - 1][1] + 1) / 2;
        ++nl;
    }
    return nl;
}

int mg_bottom_level(int nl, const int (*dims)[2]) {
    // the first level from which the remaining levels' five planes fit in the bottom kernel's LDS
    for (int l = 0; l < nl; ++l) {
        long long tot = 0;
        for (int m = l; m < nl; ++m) tot += (long long)kMgPlanes * dims[m][0] * dims[m][1];
        if (tot <= kMgBottomFloats && (long long)dims[l][0] * dims[l][1] <= kMgBottomCells) return l;
    }
    return -1;   // the coarsest level alone does not fit (a very elongated grid)
}

namespace {

// Tiles of a level above the bottom: 64x32 (512 threads) while that gives at least 256 tiles,
// else 32x16 (256 threads): on the small levels the kernels are latency bound, and smaller tiles
// spread the same work over more workgroups with less serial work in each.
inline bool mgSmallTiles(const MgLevel &F) { return ((F.nx + 63) / 64) * ((F.ny + 31) / 32) < 256; }

template <int TX, int TY, int NT, int NU>
void launchDownT(const Lc &L, const MgLaunch &G, int l, bool xz) {
    const MgLevel &F = G.lv[l];
    const dim3 grid((unsigned)(((F.nx + TX - 1) / TX) * ((F.ny + TY - 1) / TY)));
    if (l == 0) {
        if (xz) k_mg_down2d<TX, TY, NT, NU, true, true><<<grid, NT, 0, L.s>>>(*L.P, F, G.lv[1]);
        else k_mg_down2d<TX, TY, NT, NU, true, false><<<grid, NT, 0, L.s>>>(*L.P, F, G.lv[1]);
    } else {
        k_mg_down2d<TX, TY, NT, NU, false, true><<<grid, NT, 0, L.s>>>(*L.P, F, G.lv[l + 1]);
    }
}

template <int TX, int TY, int NT, int NU>
void launchUpT(const Lc &L, const MgLaunch &G, int l) {
    const MgLevel &F = G.lv[l];
    const dim3 grid((unsigned)(((F.nx + TX - 1) / TX) * ((F.ny + TY - 1) / TY)));
    if (l == 0) k_mg_up2d<TX, TY, NT, NU, true><<<grid, NT, 0, L.s>>>(*L.P, F, G.lv[1], G.over);
    else k_mg_up2d<TX, TY, NT, NU, false><<<grid, NT, 0, L.s>>>(*L.P, F, G.lv[l + 1], G.over);
}

template <int NU>
void launchDown(const Lc &L, const MgLaunch &G, int l, bool xz) {
    if (mgSmallTiles(G.lv[l])) launchDownT<32, 16, 256, NU>(L, G, l, xz);
    else launchDownT<64, 32, 512, NU>(L, G, l, xz);
}

template <int NU>
void launchUp(const Lc &L, const MgLaunch &G, int l) {
    if (mgSmallTiles(G.lv[l])) launchUpT<32, 16, 256, NU>(L, G, l);
    else launchUpT<64, 32, 512, NU>(L, G, l);
}

template <int NU>
void launchBottom(const Lc &L, const MgLaunch &G, bool xz, int recordSweeps) {
    constexpr int NT = 1024;
    if (G.lb == 0) {
        if (xz) k_mg_bottom2d<NU, true, true, NT><<<1, NT, 0, L.s>>>(*L.P, G.dev, 0, G.nl, G.ncoarse, G.over, recordSweeps);
        else k_mg_bottom2d<NU, true, false, NT><<<1, NT, 0, L.s>>>(*L.P, G.dev, 0, G.nl, G.ncoarse, G.over, recordSweeps);
    } else {
        k_mg_bottom2d<NU, false, true, NT><<<1, NT, 0, L.s>>>(*L.P, G.dev, G.lb, G.nl, G.ncoarse, G.over, recordSweeps);
    }
}

template <int NU>
void cycles(const Lc &L, const MgLaunch &G) {
    for (int c = 0; c < G.cycles; ++c) {
        for (int l = 0; l < G.lb; ++l) launchDown<NU>(L, G, l, c == 0);
        launchBottom<NU>(L, G, G.lb > 0 || c == 0, c == 0 ? G.cycles * 2 * G.nu : 0);
        for (int l = G.lb - 1; l >= 0; --l) launchUp<NU>(L, G, l);
    }
}

}  // namespace

void launch_mg_solve(const Lc &L, const MgLaunch &G) {
    // the matrices of levels 1 .. lb (the bottom kernel makes the coarser ones)
    for (int s = 0; s < G.lb; s += kMgCoarsenDepth) {
        MgCoarsen cz{};
        cz.s = G.lv[s];
        cz.m = std::min(kMgCoarsenDepth, G.lb - s);
        for (int q = 0; q < cz.m; ++q) cz.d[q] = G.lv[s + 1 + q];
        const dim3 grid((unsigned)(((cz.s.nx + 63) / 64) * ((cz.s.ny + 63) / 64)));
        if (s == 0) k_mg_coarsen2d<true><<<grid, 256, 0, L.s>>>(*L.P, cz);
        else k_mg_coarsen2d<false><<<grid, 256, 0, L.s>>>(*L.P, cz);
    }
    if (G.nu == 1) cycles<1>(L, G);
    else if (G.nu == 3) cycles<3>(L, G);
    else cycles<2>(L, G);
}

}  // namespace ifl
