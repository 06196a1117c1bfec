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
    constexpr int HX = TX + 4, HY = TY + 4;
    constexpr bool PC = KIND == COEF_PCORR;
    __shared__ float sphi[NEQ][HY][HX];
    __shared__ float srinv[PC ? HY : 1][PC ? HX : 1];
    if (caseDone(P)) return;
    const int nx = P.nx, ny = P.ny;
    const int tilesX = (nx + TX - 1) / TX;
    const int by = blockIdx.x / tilesX;
    const int bx = blockIdx.x - by * tilesX;
    const int x0 = bx * TX, y0 = by * TY;
    const int nblk = gridDim.x;
    const SweepEq eqs[2] = {e0, e1};
    bool act[2] = {false, false};
#pragma unroll
    for (int q = 0; q < NEQ; ++q) act[q] = sweepRuns(P, eqs[q].slot, sweep, nblk, eqs[q].tol);

    // stage the old field (and 1/aC) on tile + 2-cell halo; outside the grid the value is 0
    for (int e = threadIdx.x; e < HX * HY; e += kBlock) {
        const int ly = e / HX, lx = e - ly * HX;
        const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
        const bool in = gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        const int g = gy * nx + gx;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) sphi[q][ly][lx] = in ? eqs[q].in[g] : 0.0f;
        if constexpr (PC) srinv[ly][lx] = in ? coef[g] : 0.0f;
    }
    __syncthreads();

    const double Kx = -0.5 * P.rho * P.Ax * P.Ax;   // kernels.cuh:518 prefactor, same evaluation order
    const double Ky = -0.5 * P.rho * P.Ay * P.Ay;
    double acc[2] = {0.0, 0.0};

    auto relax = [&](int gx, int gy) {
        const int lx = gx - x0 + 2, ly = gy - y0 + 2;
        const int g = gy * nx + gx;
        float a[7];
        if constexpr (KIND == COEF_STORED) {
#pragma unroll
            for (int s = 0; s < 5; ++s) a[s] = coef[g + P.N * s];
        } else if constexpr (KIND == COEF_MOMENTUM) {
            const float ue = P.uf[ufi(P, gx + 1, gy, 0)];
            const float uw = P.uf[ufi(P, gx, gy, 0)];
            const float vn = P.vf[vfi(P, gx, gy + 1, 0)];
            const float vs = P.vf[vfi(P, gx, gy, 0)];
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
#pragma unroll
        for (int q = 0; q < NEQ; ++q) {
            if (!act[q]) continue;
            const float phi_C = sphi[q][ly][lx];
            float newPhi = eqs[q].src[g] - a[sE] * sphi[q][ly][lx + 1] - a[sW] * sphi[q][ly][lx - 1]
                           - a[sN] * sphi[q][ly + 1][lx] - a[sS] * sphi[q][ly - 1][lx];
            newPhi /= a[sC];
            const float d = eqs[q].rlx * (newPhi - phi_C);
            sphi[q][ly][lx] = phi_C + d;
            if (owned) acc[q] += (double)d * (double)d;
        }
    };

    if (act[0] || act[1]) {
        // red cells ((gx+gy) even) of the tile and its 1-cell ring
        constexpr int RW = TX / 2 + 1;
        for (int e = threadIdx.x; e < (TY + 2) * RW; e += kBlock) {
            const int r = e / RW, kk = e - r * RW;
            const int gy = y0 - 1 + r;
            const int gx = x0 - 1 + 2 * kk + ((x0 - 1 + gy) & 1);
            if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) relax(gx, gy);
        }
        __syncthreads();
        // black cells of the tile
        constexpr int BW = TX / 2;
        for (int e = threadIdx.x; e < TY * BW; e += kBlock) {
            const int r = e / BW, kk = e - r * BW;
            const int gy = y0 + r;
            const int gx = x0 + 2 * kk + ((x0 + gy + 1) & 1);
            if (gx < nx && gy < ny) relax(gx, gy);
        }
        __syncthreads();
    }
    // write the tile (a stopped equation is copied through unchanged, keeping the ping-pong parity)
    for (int e = threadIdx.x; e < TX * TY; e += kBlock) {
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
