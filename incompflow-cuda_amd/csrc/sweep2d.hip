// sweep2d.hip - K (1 or 2) full red-black Gauss-Seidel sweeps per launch for 2D grids.
//
// A workgroup owns a TX x TY tile.  It stages the old field on the tile plus a 2K-cell halo in
// LDS and runs the 2K half sweeps R1 B1 (R2 B2) on shrinking regions: half sweep p updates the
// cells of its colour within 2K-1-p cells of the tile, which is exactly the set the later half
// sweeps read.  Cells outside the tile are recomputed from the same old values their owning tile
// uses, so they are bit-identical, and each launch equals the reference's red pass followed by
// its black pass, K times (GaussSeidelIterateKernel, kernels.cuh:926-1005, in red-black order).
// Reads come from one buffer and writes go to the other (ping-pong), so no workgroup reads a
// value another workgroup has already updated.
//
// Coefficients are not read from stored planes in the hot loop:
//   * momentum: recomputed from the four face velocities and the boundary rows (the same
//     functions as the assembly kernel, ifl_stencil.hpp), one set for u and v, which are swept
//     together in one launch since they share the matrix;
//   * pressure correction: recomputed from 1/aC of the momentum matrix (calcPresCorrLinkCoef-
//     Kernel, kernels.cuh:517-529).
// A p' sweep moves 4 words per cell (field in/out, 1/aC, source) instead of the 8 of a stored
// 5-point stencil; with K = 2 the field, 1/aC and source are read once for two sweeps.
// Every global load of the tile is issued before the stop decision so their latency overlaps.
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_stencil.hpp"

namespace ifl {

template <int TX, int TY, int K, int NEQ, int KIND, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_rb_tile2d(CaseParams P0, BatchMap M, SweepSel e0, SweepSel e1, int launch) {
    constexpr int H = 2 * K;                                   // field halo
    constexpr int R0 = 2 * K - 1;                              // ring of the first half sweep
    constexpr int HX = TX + 2 * H, HY = TY + 2 * H;
    constexpr int SX = TX + 2 * R0, SY = TY + 2 * R0;          // cells ever updated (sources)
    constexpr bool PC = KIND == COEF_PCORR, MOM = KIND == COEF_MOMENTUM;
    constexpr int UX = SX + 1, UY = SY;                        // uf faces of the updated cells
    constexpr int VX = SX, VY = SY + 1;                        // vf faces of the updated cells
    constexpr int NST = (HX * HY + kBlock - 1) / kBlock;
    constexpr int NSR = (SX * SY + kBlock - 1) / kBlock;
    constexpr int NUT = MOM ? (UX * UY + kBlock - 1) / kBlock : 1;
    constexpr int NVT = MOM ? (VX * VY + kBlock - 1) / kBlock : 1;
    __shared__ float sphi[NEQ][HY][HX];
    __shared__ float ssrc[NEQ][SY][SX];
    __shared__ float srinv[PC ? HY : 1][PC ? HX : 1];
    __shared__ float suf[MOM ? UY : 1][MOM ? UX : 1];
    __shared__ float svf[MOM ? VY : 1][MOM ? VX : 1];
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_TILE, blk);
    if (caseDone(P)) return;
    const int nx = P.nx, ny = P.ny;
    const int tilesX = (nx + TX - 1) / TX;
    const int by = blk.flat / tilesX;
    const int bx = blk.flat - by * tilesX;
    const int x0 = bx * TX, y0 = by * TY;
    const SweepSel eqs[2] = {e0, e1};
    const float *in[2], *src[2];
    float *out[2];
    float rlx[2];
#pragma unroll
    for (int q = 0; q < NEQ; ++q) {
        float *fld = slotField(P, eqs[q].slot), *alt = P.alt[eqs[q].slot];
        in[q] = (launch & 1) ? alt : fld;
        out[q] = (launch & 1) ? fld : alt;
        src[q] = slotSource(P, eqs[q].slot);
        rlx[q] = P.relax[eqs[q].q];
    }
    const float *coef = PC ? P.rinv : slotCoef(P, eqs[0].slot);
    const int tid = threadIdx.x;

    // ---- every global load of the tile first ----
    float rphi[NEQ][NST], rr[PC ? NST : 1], rsrc[NEQ][NSR];
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int e = tid + m * kBlock;
        const int ly = e / HX, lx = e - ly * HX;
        const int gx = x0 - H + lx, gy = y0 - H + ly;
        const bool ok = e < HX * HY && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        const int g = gy * nx + gx;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) rphi[q][m] = ok ? in[q][g] : 0.0f;
        if constexpr (PC) rr[m] = ok ? coef[g] : 0.0f;
    }
#pragma unroll
    for (int m = 0; m < NSR; ++m) {
        const int e = tid + m * kBlock;
        const int ly = e / SX, lx = e - ly * SX;
        const int gx = x0 - R0 + lx, gy = y0 - R0 + ly;
        const bool ok = e < SX * SY && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) rsrc[q][m] = ok ? src[q][gy * nx + gx] : 0.0f;
    }
    float ruf[NUT], rvf[NVT];
    if constexpr (MOM) {
#pragma unroll
        for (int m = 0; m < NUT; ++m) {
            const int e = tid + m * kBlock;
            const int ly = e / UX, lx = e - ly * UX;
            const int fx = x0 - R0 + lx, gy = y0 - R0 + ly;
            const bool in = e < UX * UY && fx >= 0 && fx <= nx && gy >= 0 && gy < ny;
            ruf[m] = in ? P.uf[ufi(P, fx, gy, 0)] : 0.0f;
        }
#pragma unroll
        for (int m = 0; m < NVT; ++m) {
            const int e = tid + m * kBlock;
            const int ly = e / VX, lx = e - ly * VX;
            const int gx = x0 - R0 + lx, fy = y0 - R0 + ly;
            const bool in = e < VX * VY && gx >= 0 && gx < nx && fy >= 0 && fy <= ny;
            rvf[m] = in ? P.vf[vfi(P, gx, fy, 0)] : 0.0f;
        }
    }

    // ---- stop decision of this launch from the previous launch's partials ----
    bool act[2] = {false, false};
    int kq[2] = {0, 0};
#pragma unroll
    for (int q = 0; q < NEQ; ++q) {
        act[q] = eqs[q].k > 0 && launchRuns(P, blk, eqs[q].slot, launch, eqs[q].kPrev, blk.n, P.tol[eqs[q].q]);
        kq[q] = act[q] ? eqs[q].k : 0;
    }

    // ---- stage to LDS ----
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int e = tid + m * kBlock;
        if (e < HX * HY) {
            const int ly = e / HX, lx = e - ly * HX;
#pragma unroll
            for (int q = 0; q < NEQ; ++q) sphi[q][ly][lx] = rphi[q][m];
            if constexpr (PC) srinv[ly][lx] = rr[m];
        }
    }
#pragma unroll
    for (int m = 0; m < NSR; ++m) {
        const int e = tid + m * kBlock;
        if (e < SX * SY) {
            const int ly = e / SX, lx = e - ly * SX;
#pragma unroll
            for (int q = 0; q < NEQ; ++q) ssrc[q][ly][lx] = rsrc[q][m];
        }
    }
    if constexpr (MOM) {
#pragma unroll
        for (int m = 0; m < NUT; ++m) {
            const int e = tid + m * kBlock;
            if (e < UX * UY) suf[e / UX][e % UX] = ruf[m];
        }
#pragma unroll
        for (int m = 0; m < NVT; ++m) {
            const int e = tid + m * kBlock;
            if (e < VX * VY) svf[e / VX][e % VX] = rvf[m];
        }
    }
    __syncthreads();

    const double Kx = -0.5 * P.rho * P.Ax * P.Ax;   // kernels.cuh:518 prefactor, same evaluation order
    const double Ky = -0.5 * P.rho * P.Ay * P.Ay;
    double acc[K][2];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j][0] = acc[j][1] = 0.0;

    // relax the cell (gx, gy) in half sweep p (sweep p/2 of this launch)
    auto relax = [&](int gx, int gy, int p) {
        const int lx = gx - x0 + H, ly = gy - y0 + H;     // field / 1/aC position
        const int sx = gx - x0 + R0, sy = gy - y0 + R0;   // source / face position
        float a[7];
        if constexpr (KIND == COEF_STORED) {
            const int g = gy * nx + gx;
#pragma unroll
            for (int s = 0; s < 5; ++s) a[s] = coef[g + P.N * s];
        } else if constexpr (MOM) {
            const float uw = suf[sy][sx], ue = suf[sy][sx + 1];
            const float vs = svf[sy][sx], vn = svf[sy + 1][sx];
            momLinkCoef<2>(P, ue, uw, vn, vs, 0.0f, 0.0f, a);
            momBcRowsCoef<2>(P, gx, gy, 0, a);
        } else {
            const float rC = srinv[ly][lx];
            a[sE] = (gx < nx - 1) ? (float)(Kx * (rC + srinv[ly][lx + 1])) : 0.0f;
            a[sW] = (gx > 0) ? (float)(Kx * (rC + srinv[ly][lx - 1])) : 0.0f;
            a[sN] = (gy < ny - 1) ? (float)(Ky * (rC + srinv[ly + 1][lx])) : 0.0f;
            a[sS] = (gy > 0) ? (float)(Ky * (rC + srinv[ly - 1][lx])) : 0.0f;
            a[sC] = -(a[sE] + a[sW] + a[sN] + a[sS]);
        }
        const bool owned = gx >= x0 && gx < x0 + TX && gy >= y0 && gy < y0 + TY;
        const int j = p >> 1;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) {
            if (2 * kq[q] <= p) continue;
            const float phi_C = sphi[q][ly][lx];
            float newPhi = ssrc[q][sy][sx] - a[sE] * sphi[q][ly][lx + 1] - a[sW] * sphi[q][ly][lx - 1]
                           - a[sN] * sphi[q][ly + 1][lx] - a[sS] * sphi[q][ly - 1][lx];
            newPhi /= a[sC];
            const float d = rlx[q] * (newPhi - phi_C);
            sphi[q][ly][lx] = phi_C + d;
            if (owned) {
#pragma unroll
                for (int jj = 0; jj < K; ++jj)
                    if (jj == j) acc[jj][q] += (double)d * (double)d;
            }
        }
    };

    const int nph = 2 * (kq[0] > kq[NEQ - 1] ? kq[0] : kq[NEQ - 1]);
#pragma unroll
    for (int p = 0; p < 2 * K; ++p) {
        if (p >= nph) break;
        const int color = p & 1, ring = 2 * K - 1 - p;
        const int xs = x0 - ring, ys = y0 - ring;
        const int half = (TX + 2 * ring) / 2, rows = TY + 2 * ring;
        for (int e = tid; e < rows * half; e += kBlock) {
            const int r = e / half, kk = e - r * half;
            const int gy = ys + r;
            const int gx = xs + 2 * kk + ((xs + gy + color) & 1);
            if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) relax(gx, gy, p);
        }
        __syncthreads();
    }
    // write the tile (a stopped equation is copied through unchanged, keeping the ping-pong parity)
    for (int e = tid; e < TX * TY; e += kBlock) {
        const int r = e / TX, xx = e - r * TX;
        const int gx = x0 + xx, gy = y0 + r;
        if (gx < nx && gy < ny) {
            const int g = gy * nx + gx;
#pragma unroll
            for (int q = 0; q < NEQ; ++q) out[q][g] = sphi[q][r + H][xx + H];
        }
    }
#pragma unroll
    for (int q = 0; q < NEQ; ++q)
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j < kq[q]) {
                double a1[1] = {acc[j][q]};
                storeBlockSums<1>(a1, sweepPart(P, eqs[q].slot, launch, j) + blk.flat);
            }
}

template <int TX, int TY, int K, bool BATCH>
static void launchTile(const Lc &L, int kind, const SweepSel &e0, const SweepSel &e1, int neq, int launch) {
    const CaseParams &P = *L.P;
    const dim3 grid = L.batch ? dim3((unsigned)L.total[MAP_TILE])
                              : dim3((unsigned)(((P.nx + TX - 1) / TX) * ((P.ny + TY - 1) / TY)));
    if (neq == 2) {
        if (kind == COEF_MOMENTUM) k_rb_tile2d<TX, TY, K, 2, COEF_MOMENTUM, BATCH><<<grid, kBlock, 0, L.s>>>(P, L.M, e0, e1, launch);
        else k_rb_tile2d<TX, TY, K, 2, COEF_STORED, BATCH><<<grid, kBlock, 0, L.s>>>(P, L.M, e0, e1, launch);
    } else {
        if (kind == COEF_MOMENTUM) k_rb_tile2d<TX, TY, K, 1, COEF_MOMENTUM, BATCH><<<grid, kBlock, 0, L.s>>>(P, L.M, e0, e1, launch);
        else if (kind == COEF_PCORR) k_rb_tile2d<TX, TY, K, 1, COEF_PCORR, BATCH><<<grid, kBlock, 0, L.s>>>(P, L.M, e0, e1, launch);
        else k_rb_tile2d<TX, TY, K, 1, COEF_STORED, BATCH><<<grid, kBlock, 0, L.s>>>(P, L.M, e0, e1, launch);
    }
}

template <int K, bool BATCH>
static void launchK(const Lc &L, int kind, const SweepSel &e0, const SweepSel &e1, int neq, int launch) {
    switch (L.tile) {
        case TILE_64x16: launchTile<64, 16, K, BATCH>(L, kind, e0, e1, neq, launch); break;
        case TILE_64x8: launchTile<64, 8, K, BATCH>(L, kind, e0, e1, neq, launch); break;
        case TILE_32x8: launchTile<32, 8, K, BATCH>(L, kind, e0, e1, neq, launch); break;
        default: launchTile<32, 4, K, BATCH>(L, kind, e0, e1, neq, launch); break;
    }
}

static int tileX(int tile) { return (tile == TILE_64x16 || tile == TILE_64x8) ? 64 : 32; }
static int tileY(int tile) { return tile == TILE_64x16 ? 16 : tile == TILE_32x4 ? 4 : 8; }

int tile_dimx(int nx, int tile) { return (nx + tileX(tile) - 1) / tileX(tile); }
int tile_count(int nx, int ny, int tile) { return tile_dimx(nx, tile) * ((ny + tileY(tile) - 1) / tileY(tile)); }

int default_tile(int nx, int ny) {
    if (tile_count(nx, ny, TILE_64x16) >= 768) return TILE_64x16;
    if (tile_count(nx, ny, TILE_64x8) >= 512) return TILE_64x8;
    if (tile_count(nx, ny, TILE_32x8) >= 256) return TILE_32x8;
    return TILE_32x4;
}

void launch_rb_sweep2d(const Lc &L, int kind, const SweepSel &e0, const SweepSel &e1, int neq, int launch, int block) {
    if (L.batch) {
        if (block == 2) launchK<2, true>(L, kind, e0, e1, neq, launch);
        else launchK<1, true>(L, kind, e0, e1, neq, launch);
    } else {
        if (block == 2) launchK<2, false>(L, kind, e0, e1, neq, launch);
        else launchK<1, false>(L, kind, e0, e1, neq, launch);
    }
}

}  // namespace ifl
