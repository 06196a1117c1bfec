// mg2d.hip - gfx950 kernels of the multigrid p' solve on 2D grids (IFL_PSOLVE_MG).
//
// Per outer iteration: k_mg_coarsen2d builds the coarse matrices (one pass makes six levels from
// 64x64 tiles of its source level, each tile summing its own cells hierarchically in LDS; the
// couplings of a tile's west and south edge cells towards -x and -y come from the same sums over
// the neighbouring column and row).  Per V-cycle: k_mg_down2d on every level above the bottom
// (NU red-black sweeps from x = 0 on a 64x32 tile staged with a (2NU+1)-cell halo, then the FP64
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
__global__ __launch_bounds__(256) void k_mg_coarsen2d(CaseParams P, MgCoarsen G, int sweepsPerSolve) {
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
    if (L0 && t == 0 && tid == 0) {     // the p' solve's sweep count (fine-level sweeps)
        P.st->sm[3].sweeps = sweepsPerSolve;
        P.st->sm[3].lastSum = 0.0;
        P.st->sweepsIter[3] += sweepsPerSolve;
    }
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
template <int NU, bool L0, bool XZ>
__global__ __launch_bounds__(512) void k_mg_down2d(CaseParams P, MgLevel F, MgLevel Cc) {
    constexpr int TX = 64, TY = 32, NT = 512, H = 2 * NU + 1;
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
template <int NU, bool L0>
__global__ __launch_bounds__(512) void k_mg_up2d(CaseParams P, MgLevel F, MgLevel Cc, float over) {
    constexpr int TX = 64, TY = 32, NT = 512, H = 2 * NU;
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
constexpr int kMgBottomFloats = 5 * 5600;   // x, b, C, E, N of the bottom levels (row major)

template <int NT>
__device__ __forceinline__ void mgSmoothLds(float *L, int nx, int ny, int sweeps) {
    const int n = nx * ny;
    float *x = L;
    const float *b = L + n, *C = L + 2 * n, *E = L + 3 * n, *Nn = L + 4 * n;
    for (int s = 0; s < sweeps; ++s)
        for (int col = 0; col < 2; ++col) {
            for (int c = threadIdx.x; c < n; c += NT) {
                const int j = c / nx, i = c - j * nx;
                if (((i + j) & 1) != col) continue;
                const float xE = i < nx - 1 ? x[c + 1] : 0.0f, xW = i > 0 ? x[c - 1] : 0.0f;
                const float xN = j < ny - 1 ? x[c + nx] : 0.0f, xS = j > 0 ? x[c - nx] : 0.0f;
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

template <int NU, bool L0, bool XZ, int NT>
__global__ __launch_bounds__(NT) void k_mg_bottom2d(CaseParams P, const MgLevel *lv, int lb, int nl, int ncoarse,
                                                    float over) {
    __shared__ float lds[kMgBottomFloats];
    if (caseDone(P)) return;
    const int tid = threadIdx.x;
    // load: level lb (x, b, matrix), the coarser levels' matrices (x = 0)
    int base = 0;
    for (int l = lb; l < nl; ++l) {
        const int nx = lv[l].nx, ny = lv[l].ny, n = nx * ny;
        float *L = lds + base;
        for (int c = tid; c < n; c += NT) {
            float xv = 0.0f, bv = 0.0f, cv, ev, nv;
            if (L0 && l == 0) {
                if (!XZ) xv = P.pc[c];
                bv = P.bp[c];
                ev = P.pcoef[c + (size_t)P.N * sE];
                nv = P.pcoef[c + (size_t)P.N * sN];
                cv = P.pcoef[c] + mgOutlet0(P, c % nx, c / nx, c);
            } else {
                if (l == lb) {
                    if (!XZ) xv = lv[l].x[c];
                    bv = lv[l].b[c];
                }
                cv = lv[l].C[c];
                ev = lv[l].E[c];
                nv = lv[l].N[c];
            }
            L[c] = xv;
            L[n + c] = bv;
            L[2 * n + c] = cv;
            L[3 * n + c] = ev;
            L[4 * n + c] = nv;
        }
        base += 5 * n;
    }
    __syncthreads();
    // down
    base = 0;
    for (int l = lb; l < nl - 1; ++l) {
        const int nx = lv[l].nx, ny = lv[l].ny, n = nx * ny;
        const int cnx = lv[l + 1].nx, cny = lv[l + 1].ny, cn = cnx * cny;
        float *L = lds + base, *Lc = lds + base + 5 * n;
        mgSmoothLds<NT>(L, nx, ny, NU);
        for (int C = tid; C < cn; C += NT) {
            const int J = C / cnx, I = C - J * cnx;
            const bool hasI1 = 2 * I + 1 < nx, hasJ1 = 2 * J + 1 < ny;
            double acc = 0.0;
            acc += mgResidualLds(L, nx, ny, 2 * I, 2 * J);
            if (hasI1) acc += mgResidualLds(L, nx, ny, 2 * I + 1, 2 * J);
            if (hasJ1) {
                acc += mgResidualLds(L, nx, ny, 2 * I, 2 * J + 1);
                if (hasI1) acc += mgResidualLds(L, nx, ny, 2 * I + 1, 2 * J + 1);
            }
            Lc[cn + C] = (float)acc;
            Lc[C] = 0.0f;
        }
        __syncthreads();
        base += 5 * n;
    }
    {
        const int nx = lv[nl - 1].nx, ny = lv[nl - 1].ny;
        mgSmoothLds<NT>(lds + base, nx, ny, ncoarse);
    }
    // up
    for (int l = nl - 2; l >= lb; --l) {
        const int nx = lv[l].nx, ny = lv[l].ny, n = nx * ny, cnx = lv[l + 1].nx;
        float *Lc = lds + base;
        base -= 5 * n;
        float *L = lds + base;
        for (int c = tid; c < n; c += NT) {
            const int j = c / nx, i = c - j * nx;
            L[c] = L[c] + over * Lc[(j >> 1) * cnx + (i >> 1)];
        }
        __syncthreads();
        mgSmoothLds<NT>(L, nx, ny, NU);
    }
    {
        const int n = lv[lb].nx * lv[lb].ny;
        float *xg = (L0 && lb == 0) ? P.pc : lv[lb].xu;      // what the up pass reads as parents
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
        for (int m = l; m < nl; ++m) tot += 5LL * dims[m][0] * dims[m][1];
        if (tot <= kMgBottomFloats && (long long)dims[l][0] * dims[l][1] <= kMgBottomCells) return l;
    }
    return -1;   // the coarsest level alone does not fit (a very elongated grid)
}

namespace {

template <int NU>
void launchDown(const Lc &L, const MgLaunch &G, int l, bool xz) {
    const MgLevel &F = G.lv[l];
    const dim3 grid((unsigned)(((F.nx + 63) / 64) * ((F.ny + 31) / 32)));
    if (l == 0) {
        if (xz) k_mg_down2d<NU, true, true><<<grid, 512, 0, L.s>>>(*L.P, F, G.lv[1]);
        else k_mg_down2d<NU, true, false><<<grid, 512, 0, L.s>>>(*L.P, F, G.lv[1]);
    } else {
        k_mg_down2d<NU, false, true><<<grid, 512, 0, L.s>>>(*L.P, F, G.lv[l + 1]);
    }
}

template <int NU>
void launchUp(const Lc &L, const MgLaunch &G, int l) {
    const MgLevel &F = G.lv[l];
    const dim3 grid((unsigned)(((F.nx + 63) / 64) * ((F.ny + 31) / 32)));
    if (l == 0) k_mg_up2d<NU, true><<<grid, 512, 0, L.s>>>(*L.P, F, G.lv[1], G.over);
    else k_mg_up2d<NU, false><<<grid, 512, 0, L.s>>>(*L.P, F, G.lv[l + 1], G.over);
}

template <int NU>
void launchBottom(const Lc &L, const MgLaunch &G, bool xz) {
    constexpr int NT = 512;
    if (G.lb == 0) {
        if (xz) k_mg_bottom2d<NU, true, true, NT><<<1, NT, 0, L.s>>>(*L.P, G.dev, 0, G.nl, G.ncoarse, G.over);
        else k_mg_bottom2d<NU, true, false, NT><<<1, NT, 0, L.s>>>(*L.P, G.dev, 0, G.nl, G.ncoarse, G.over);
    } else {
        k_mg_bottom2d<NU, false, true, NT><<<1, NT, 0, L.s>>>(*L.P, G.dev, G.lb, G.nl, G.ncoarse, G.over);
    }
}

template <int NU>
void cycles(const Lc &L, const MgLaunch &G) {
    for (int c = 0; c < G.cycles; ++c) {
        for (int l = 0; l < G.lb; ++l) launchDown<NU>(L, G, l, c == 0);
        launchBottom<NU>(L, G, G.lb > 0 || c == 0);
        for (int l = G.lb - 1; l >= 0; --l) launchUp<NU>(L, G, l);
    }
}

}  // namespace

void launch_mg_solve(const Lc &L, const MgLaunch &G) {
    const int sweeps = G.cycles * 2 * G.nu;
    for (int s = 0; s < G.nl - 1; s += kMgCoarsenDepth) {
        MgCoarsen cz{};
        cz.s = G.lv[s];
        cz.m = std::min(kMgCoarsenDepth, G.nl - 1 - s);
        for (int q = 0; q < cz.m; ++q) cz.d[q] = G.lv[s + 1 + q];
        const dim3 grid((unsigned)(((cz.s.nx + 63) / 64) * ((cz.s.ny + 63) / 64)));
        if (s == 0) k_mg_coarsen2d<true><<<grid, 256, 0, L.s>>>(*L.P, cz, sweeps);
        else k_mg_coarsen2d<false><<<grid, 256, 0, L.s>>>(*L.P, cz, sweeps);
    }
    if (G.nu == 1) cycles<1>(L, G);
    else if (G.nu == 3) cycles<3>(L, G);
    else cycles<2>(L, G);
}

}  // namespace ifl
