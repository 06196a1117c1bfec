// sweep2d.hip - one full red-black Gauss-Seidel sweep per launch for 2D grids.
//
// A workgroup owns a TX x TY tile.  It stages the old field on the tile plus a 2-cell halo in
// LDS, relaxes the red cells of the tile and of a 1-cell ring around it (the ring's red values
// are recomputed from the same old values their owning tile uses, so they are bit-identical),
// then relaxes the tile's black cells against those red values and writes the tile.  Reads come
// from one buffer and writes go to the other (ping-pong), so no workgroup ever reads a value
// another workgroup has already updated: one launch equals the reference's red pass followed by
// its black pass, bit for bit (GaussSeidelIterateKernel, kernels.cuh:926-1005, in red-black order).
//
// Coefficients are not read from the stored planes in the hot loop:
//   * momentum: recomputed from the four face velocities and the boundary rows (the same
//     functions as the assembly kernel, ifl_stencil.hpp), one set for u and v, which are swept
//     together in one launch since they share the matrix;
//   * pressure correction: recomputed from 1/aC of the momentum matrix (calcPresCorrLinkCoef-
//     Kernel, kernels.cuh:517-529), staged in LDS with the field.
// This moves 4 words per cell and equation per sweep instead of 8.
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_stencil.hpp"

namespace ifl {

template <int TX, int TY, int NEQ, int KIND>
__global__ __launch_bounds__(kBlock) void k_rb_tile2d(CaseParams P, const float *coef, SweepEq e0, SweepEq e1,
                                                      int sweep) {
    constexpr int HX = TX + 4, HY = TY + 4;                    // field + 2-cell halo
    constexpr int NST = (HX * HY + kBlock - 1) / kBlock;       // staged values per thread
    constexpr int RW = TX / 2 + 1, NR = (TY + 2) * RW;         // red cells of tile + ring
    constexpr int NRT = (NR + kBlock - 1) / kBlock;
    constexpr int BW = TX / 2, NB = TY * BW;                   // black cells of the tile
    constexpr int NBT = (NB + kBlock - 1) / kBlock;
    constexpr bool PC = KIND == COEF_PCORR, MOM = KIND == COEF_MOMENTUM;
    constexpr int UX = TX + 3, UY = TY + 2;                    // uf faces x0-1..x0+TX+1, rows y0-1..y0+TY
    constexpr int VX = TX + 2, VY = TY + 3;                    // vf faces of cells x0-1..x0+TX, rows y0-1..y0+TY+1
    constexpr int NUT = MOM ? (UX * UY + kBlock - 1) / kBlock : 1;
    constexpr int NVT = MOM ? (VX * VY + kBlock - 1) / kBlock : 1;
    __shared__ float sphi[NEQ][HY][HX];
    __shared__ float srinv[PC ? HY : 1][PC ? HX : 1];
    __shared__ float suf[MOM ? UY : 1][MOM ? UX : 1];
    __shared__ float svf[MOM ? VY : 1][MOM ? VX : 1];
    if (caseDone(P)) return;
    const int nx = P.nx, ny = P.ny;
    const int tilesX = (nx + TX - 1) / TX;
    const int by = blockIdx.x / tilesX;
    const int bx = blockIdx.x - by * tilesX;
    const int x0 = bx * TX, y0 = by * TY;
    const SweepEq eqs[2] = {e0, e1};
    const int tid = threadIdx.x;

    // ---- issue every global load of the tile first (latency overlaps the stop decision) ----
    float rphi[NEQ][NST], rr[PC ? NST : 1];
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int e = tid + m * kBlock;
        const int ly = e / HX, lx = e - ly * HX;
        const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
        const bool in = e < HX * HY && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        const int g = gy * nx + gx;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) rphi[q][m] = in ? eqs[q].in[g] : 0.0f;
        if constexpr (PC) rr[m] = in ? coef[g] : 0.0f;
    }
    float ruf[NUT], rvf[NVT];
    if constexpr (MOM) {
#pragma unroll
        for (int m = 0; m < NUT; ++m) {
            const int e = tid + m * kBlock;
            const int ly = e / UX, lx = e - ly * UX;
            const int fx = x0 - 1 + lx, gy = y0 - 1 + ly;
            const bool in = e < UX * UY && fx >= 0 && fx <= nx && gy >= 0 && gy < ny;
            ruf[m] = in ? P.uf[ufi(P, fx, gy, 0)] : 0.0f;
        }
#pragma unroll
        for (int m = 0; m < NVT; ++m) {
            const int e = tid + m * kBlock;
            const int ly = e / VX, lx = e - ly * VX;
            const int gx = x0 - 1 + lx, fy = y0 - 1 + ly;
            const bool in = e < VX * VY && gx >= 0 && gx < nx && fy >= 0 && fy <= ny;
            rvf[m] = in ? P.vf[vfi(P, gx, fy, 0)] : 0.0f;
        }
    }
    // red cells ((gx+gy) even) of the tile and its 1-cell ring, black cells of the tile
    int redG[NRT], blkG[NBT];
    float srcR[NEQ][NRT], srcB[NEQ][NBT];
#pragma unroll
    for (int m = 0; m < NRT; ++m) {
        const int e = tid + m * kBlock;
        const int r = e / RW, kk = e - r * RW;
        const int gy = y0 - 1 + r;
        const int gx = x0 - 1 + 2 * kk + ((x0 - 1 + gy) & 1);
        const bool in = e < NR && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        redG[m] = in ? ((gy - y0 + 2) << 16) | (gx - x0 + 2) : -1;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) srcR[q][m] = in ? eqs[q].src[gy * nx + gx] : 0.0f;
    }
#pragma unroll
    for (int m = 0; m < NBT; ++m) {
        const int e = tid + m * kBlock;
        const int r = e / BW, kk = e - r * BW;
        const int gy = y0 + r;
        const int gx = x0 + 2 * kk + ((x0 + gy + 1) & 1);
        const bool in = e < NB && gx < nx && gy < ny;
        blkG[m] = in ? ((gy - y0 + 2) << 16) | (gx - x0 + 2) : -1;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) srcB[q][m] = in ? eqs[q].src[gy * nx + gx] : 0.0f;
    }

    // ---- stop decision of this sweep from the previous sweep's partials ----
    bool act[2] = {false, false};
#pragma unroll
    for (int q = 0; q < NEQ; ++q) act[q] = sweepRuns(P, eqs[q].slot, sweep, gridDim.x, eqs[q].tol);

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
    double acc[2] = {0.0, 0.0};

    // relax cell at LDS position (ly, lx) with sources src[q]
    auto relax = [&](int ly, int lx, const float (&src)[2]) {
        const int gx = x0 + lx - 2, gy = y0 + ly - 2;
        float a[7];
        if constexpr (KIND == COEF_STORED) {
            const int g = gy * nx + gx;
#pragma unroll
            for (int s = 0; s < 5; ++s) a[s] = coef[g + P.N * s];
        } else if constexpr (MOM) {
            // uf face of cell (gx,gy) west side sits at suf[ly-1][lx-1], vf south at svf[ly-1][lx-1]
            const float uw = suf[ly - 1][lx - 1], ue = suf[ly - 1][lx];
            const float vs = svf[ly - 1][lx - 1], vn = svf[ly][lx - 1];
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
        const bool owned = lx >= 2 && lx < TX + 2 && ly >= 2 && ly < TY + 2;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) {
            if (!act[q]) continue;
            const float phi_C = sphi[q][ly][lx];
            float newPhi = src[q] - a[sE] * sphi[q][ly][lx + 1] - a[sW] * sphi[q][ly][lx - 1]
                           - a[sN] * sphi[q][ly + 1][lx] - a[sS] * sphi[q][ly - 1][lx];
            newPhi /= a[sC];
            const float d = eqs[q].rlx * (newPhi - phi_C);
            sphi[q][ly][lx] = phi_C + d;
            if (owned) acc[q] += (double)d * (double)d;
        }
    };

    if (act[0] || act[1]) {
#pragma unroll
        for (int m = 0; m < NRT; ++m)
            if (redG[m] >= 0) {
                const float src[2] = {srcR[0][m], srcR[NEQ - 1][m]};
                relax(redG[m] >> 16, redG[m] & 0xffff, src);
            }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NBT; ++m)
            if (blkG[m] >= 0) {
                const float src[2] = {srcB[0][m], srcB[NEQ - 1][m]};
                relax(blkG[m] >> 16, blkG[m] & 0xffff, src);
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
            for (int q = 0; q < NEQ; ++q) eqs[q].out[g] = sphi[q][r + 2][xx + 2];
        }
    }
#pragma unroll
    for (int q = 0; q < NEQ; ++q)
        if (act[q]) {
            double a1[1] = {acc[q]};
            storeBlockSums<1>(a1, sweepPart(P, eqs[q].slot, sweep) + blockIdx.x);
        }
}

__global__ __launch_bounds__(kBlock) void k_copy(CaseParams P, float *dst, const float *src) {
    if (caseDone(P)) return;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < P.N) dst[t] = src[t];
}

template <int TX, int TY>
static void launchTile(const CaseParams &P, int kind, const float *coef, const SweepEq &e0, const SweepEq &e1,
                       int neq, int sweep, hipStream_t s) {
    const unsigned grid = (unsigned)(((P.nx + TX - 1) / TX) * ((P.ny + TY - 1) / TY));
    if (neq == 2) {
        if (kind == COEF_MOMENTUM) k_rb_tile2d<TX, TY, 2, COEF_MOMENTUM><<<grid, kBlock, 0, s>>>(P, coef, e0, e1, sweep);
        else k_rb_tile2d<TX, TY, 2, COEF_STORED><<<grid, kBlock, 0, s>>>(P, coef, e0, e1, sweep);
    } else {
        if (kind == COEF_MOMENTUM) k_rb_tile2d<TX, TY, 1, COEF_MOMENTUM><<<grid, kBlock, 0, s>>>(P, coef, e0, e1, sweep);
        else if (kind == COEF_PCORR) k_rb_tile2d<TX, TY, 1, COEF_PCORR><<<grid, kBlock, 0, s>>>(P, coef, e0, e1, sweep);
        else k_rb_tile2d<TX, TY, 1, COEF_STORED><<<grid, kBlock, 0, s>>>(P, coef, e0, e1, sweep);
    }
}

int tile_count(int nx, int ny, int tile) {
    switch (tile) {
        case TILE_64x16: return ((nx + 63) / 64) * ((ny + 15) / 16);
        case TILE_64x8: return ((nx + 63) / 64) * ((ny + 7) / 8);
        case TILE_32x8: return ((nx + 31) / 32) * ((ny + 7) / 8);
        default: return ((nx + 31) / 32) * ((ny + 3) / 4);
    }
}

int default_tile(int nx, int ny) {
    if (tile_count(nx, ny, TILE_64x16) >= 768) return TILE_64x16;
    if (tile_count(nx, ny, TILE_64x8) >= 512) return TILE_64x8;
    if (tile_count(nx, ny, TILE_32x8) >= 256) return TILE_32x8;
    return TILE_32x4;
}

void launch_rb_sweep2d(const CaseParams &P, int kind, const float *coef, const SweepEq &e0, const SweepEq &e1, int neq,
                       int sweep, int tile, hipStream_t s) {
    switch (tile) {
        case TILE_64x16: launchTile<64, 16>(P, kind, coef, e0, e1, neq, sweep, s); break;
        case TILE_64x8: launchTile<64, 8>(P, kind, coef, e0, e1, neq, sweep, s); break;
        case TILE_32x8: launchTile<32, 8>(P, kind, coef, e0, e1, neq, sweep, s); break;
        default: launchTile<32, 4>(P, kind, coef, e0, e1, neq, sweep, s); break;
    }
}

void launch_copy(const CaseParams &P, float *dst, const float *src, hipStream_t s) {
    k_copy<<<(unsigned)((P.N + kBlock - 1) / kBlock), kBlock, 0, s>>>(P, dst, src);
}

}  // namespace ifl
