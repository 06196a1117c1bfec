// sweep3d.hip - one full red-black Gauss-Seidel sweep per launch for 3D grids.
//
// A workgroup owns a TX x TY column of the grid over ZC planes [z0, z1).  It marches up the
// column with a rolling window of four planes of the old field in LDS (2-cell halo in x and y)
// and at step t relaxes the red cells of plane t+1 (on the tile plus a 1-cell ring), then the
// black cells of plane t (on the tile), then writes plane t:
//   red(t+1) reads black values of planes t..t+2, which are still old;
//   black(t) reads red values of planes t-1..t+1, which are already new.
// The red planes z0-1 and z1 and the ring cells are recomputed from the same old values their
// owners use, so the launch equals the reference's red pass followed by its black pass bit for
// bit (GaussSeidelIterateKernel, kernels.cuh:926-1005, in red-black order, 7-point stencil).
// Reads come from one buffer and writes go to the other (ping-pong).  Coefficients are the
// stored (1+2dim) planes of the assembled matrix.
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_stencil.hpp"

namespace ifl {

constexpr int Z_TX = 64, Z_TY = 8, Z_ZC = 16;

template <int TX, int TY, int ZC, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_rb_zmarch3d(CaseParams P0, BatchMap M, SweepSel e, int launch) {
    constexpr int HX = TX + 4, HY = TY + 4, PL = HX * HY;
    constexpr int NPL = (PL + kBlock - 1) / kBlock;
    constexpr int RW = TX / 2 + 1, NR = (TY + 2) * RW;       // red cells of tile + ring per plane
    constexpr int BW = TX / 2, NB = TY * BW;                 // black cells of the tile per plane
    __shared__ float win[4][HY][HX];                         // planes z mod 4
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_ZMARCH, blk);
    if (caseDone(P)) return;
    const int nx = P.nx, ny = P.ny, nz = P.nz, N = P.N;
    const int gxc = (nx + TX - 1) / TX, gyc = (ny + TY - 1) / TY;   // column grid of the case
    const int bx = blk.flat % gxc, byz = blk.flat / gxc;
    const int by = byz % gyc, bz = byz / gyc;
    float *fld = slotField(P, e.slot), *alt = slotAlt(P, e.slot);
    const bool odd = ((launch + e.flip) & 1) != 0;
    const float *ein = odd ? alt : fld;
    float *eout = odd ? fld : alt;
    const float *esrc = slotSource(P, e.slot), *coef = slotCoef(P, e.slot);
    const float erlx = P.relax[e.q];
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * ZC;
    const int z1 = min(z0 + ZC, nz);
    const int tid = threadIdx.x;


    // issue the loads of the first four planes, then decide whether this sweep runs
    float reg[4][NPL];
    auto loadPlane = [&](int z, float (&r)[NPL]) {
#pragma unroll
        for (int m = 0; m < NPL; ++m) {
            const int q = tid + m * kBlock;
            const int ly = q / HX, lx = q - ly * HX;
            const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
            const bool in = q < PL && z >= 0 && z < nz && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
            r[m] = in ? ein[gx + nx * (gy + ny * z)] : 0.0f;
        }
    };
    auto storePlane = [&](int z, const float (&r)[NPL]) {
        float *w = &win[z & 3][0][0];
#pragma unroll
        for (int m = 0; m < NPL; ++m) {
            const int q = tid + m * kBlock;
            if (q < PL) w[q] = r[m];
        }
    };
#pragma unroll
    for (int s = 0; s < 4; ++s) loadPlane(z0 - 2 + s, reg[s]);
    const bool run = launchRuns(P, blk, e.slot, launch, e.kPrev, blk.n, P.tol[e.q]);
#pragma unroll
    for (int s = 0; s < 4; ++s) storePlane(z0 - 2 + s, reg[s]);
    __syncthreads();

    double acc = 0.0;
    auto relax = [&](int gx, int gy, int z, bool owned) {
        const int c = gx + nx * (gy + ny * z);
        const int lx = gx - x0 + 2, ly = gy - y0 + 2;
        float(*wc)[HX] = win[z & 3];
        const float phi_C = wc[ly][lx];
        const float phi_E = wc[ly][lx + 1], phi_W = wc[ly][lx - 1];
        const float phi_N = wc[ly + 1][lx], phi_S = wc[ly - 1][lx];
        const float phi_T = (z != nz - 1) ? win[(z + 1) & 3][ly][lx] : 0.0f;
        const float phi_B = (z != 0) ? win[(z - 1) & 3][ly][lx] : 0.0f;
        float newPhi = esrc[c] - coef[c + N * sE] * phi_E - coef[c + N * sW] * phi_W - coef[c + N * sN] * phi_N
                       - coef[c + N * sS] * phi_S;
        newPhi -= coef[c + N * sT] * phi_T + coef[c + N * sB] * phi_B;
        newPhi /= coef[c];
        const float d = erlx * (newPhi - phi_C);
        wc[ly][lx] = phi_C + d;
        if (owned) acc += (double)d * (double)d;
    };
    auto redPlane = [&](int z) {
        if (z < 0 || z >= nz) return;
        const bool ownPlane = z >= z0 && z < z1;
        for (int q = tid; q < NR; q += kBlock) {
            const int r = q / RW, kk = q - r * RW;
            const int gy = y0 - 1 + r;
            const int gx = x0 - 1 + 2 * kk + ((x0 - 1 + gy + z) & 1);
            if (gx >= 0 && gx < nx && gy >= 0 && gy < ny)
                relax(gx, gy, z, ownPlane && gx >= x0 && gx < x0 + TX && gy >= y0 && gy < y0 + TY);
        }
    };
    auto blackPlane = [&](int z) {
        for (int q = tid; q < NB; q += kBlock) {
            const int r = q / BW, kk = q - r * BW;
            const int gy = y0 + r;
            const int gx = x0 + 2 * kk + ((x0 + gy + z + 1) & 1);
            if (gx < nx && gy < ny) relax(gx, gy, z, true);
        }
    };

    if (run) {
        redPlane(z0 - 1);
        redPlane(z0);
        __syncthreads();
    }
    for (int t = z0; t < z1; ++t) {
        float nxt[NPL];
        loadPlane(t + 2, nxt);                 // old plane t+2 (needed by red(t+1))
        if (run) {
            // plane t+2 replaces plane t-2 (slot (t+2)&3), which no step reads any more
            storePlane(t + 2, nxt);
            __syncthreads();
            redPlane(t + 1);
            __syncthreads();
            blackPlane(t);
            __syncthreads();
        } else {
            storePlane(t + 2, nxt);
            __syncthreads();
        }
        // write plane t of the tile
        for (int q = tid; q < TX * TY; q += kBlock) {
            const int r = q / TX, xx = q - r * TX;
            const int gx = x0 + xx, gy = y0 + r;
            if (gx < nx && gy < ny) eout[gx + nx * (gy + ny * t)] = win[t & 3][r + 2][xx + 2];
        }
        __syncthreads();
    }
    if (run) {
        double a1[1] = {acc};
        storeBlockSums<1>(a1, sweepPart(P, e.slot, launch, 0) + blk.flat);
    }
}

int zmarch_blocks(const CaseParams &P) {
    return ((P.nx + Z_TX - 1) / Z_TX) * ((P.ny + Z_TY - 1) / Z_TY) * ((P.nz + Z_ZC - 1) / Z_ZC);
}

int zmarch_dimx(const CaseParams &P) { return (P.nx + Z_TX - 1) / Z_TX; }

void launch_rb_sweep3d(const Lc &L, const SweepSel &e, int launch) {
    const CaseParams &P = *L.P;
    if (L.batch) {
        k_rb_zmarch3d<Z_TX, Z_TY, Z_ZC, true><<<dim3((unsigned)L.total[MAP_ZMARCH]), kBlock, 0, L.s>>>(P, L.M, e, launch);
        return;
    }
    dim3 grid((unsigned)((P.nx + Z_TX - 1) / Z_TX), (unsigned)((P.ny + Z_TY - 1) / Z_TY),
              (unsigned)((P.nz + Z_ZC - 1) / Z_ZC));
    k_rb_zmarch3d<Z_TX, Z_TY, Z_ZC, false><<<grid, kBlock, 0, L.s>>>(P, L.M, e, launch);
}

}  // namespace ifl
