// sweep3d.hip - red-black Gauss-Seidel sweeps for 3D grids: one full sweep per launch
// (k_rb_zmarch3d, k_rb_zmarch3d_p) or, for the p' system, two (k_rb_zmarch3d_p2, further down).
//
// A workgroup owns a TX x TY column of the grid over ZC planes [z0, z1).  It marches up the
// column with a rolling window of four planes of the field in LDS (2-cell halo in x and y) and
// at step t relaxes the red cells of plane t+1 (on the tile plus a 1-cell ring), then the black
// cells of plane t (on the tile), then writes plane t:
//   red(t+1) reads black values of planes t..t+2, which are still old;
//   black(t) reads red values of planes t-1..t+1, which are already new.
// The red planes z0-1 and z1 and the ring cells are recomputed from the same old values their
// owners use, so the launch equals the reference's red pass followed by its black pass bit for
// bit (GaussSeidelIterateKernel, kernels.cuh:926-1005, in red-black order, 7-point stencil).
// Reads come from one buffer and writes go to the other (ping-pong).
//
// k_rb_zmarch3d sweeps the momentum equations (u, v, w: one matrix, the seven stored planes)
// together, NEQ of them per launch, so each coefficient is read once for all of them, and the
// temperature equation; the coefficients and sources of the cells a step relaxes are loaded
// ahead, into registers, so their latency overlaps the previous steps.  k_rb_zmarch3d_p sweeps
// the p' system, which reads only its +x, +y and +z planes: its matrix is symmetric bit for bit
// (the coefficient of a cell towards -x is the +x coefficient of its west neighbour,
// kernels.cuh:517-529 computes both from the same commutative sum) and its diagonal is
// -(E+W+N+S) - (T+B) as assembled.
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_stencil.hpp"

namespace ifl {

// Column shapes, measured per launch on the 256^3 cavity (us).  u+v+w sweep: 128x8x32 310,
// 128x8x16 334, 64x8x32 461, 64x8x16 393, 128x16x16 448, 256x4x32 522, 256x8x16 479, 128x4x32 362
// (three workgroups per CU instead of two): it moves three fields and seven coefficient planes per
// cell and prefers wide columns.  p' sweep (k_rb_zmarch3d_p,
// links staged in LDS): 32x8x32 138, 64x8x43 139, 64x16x32 139, 64x4x32 146, 32x8x64 145,
// 64x8x32 165, 64x8x64 166, 128x8x32 216 (64x8x32 with per-cell link loads: 149).
constexpr int Z_TX = 128, Z_TY = 8, Z_ZC = 32;     // momentum, temperature
constexpr int ZP_TX = 32, ZP_TY = 8, ZP_ZC = 32;   // p'

template <int TX, int TY, int ZC, int NEQ, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_rb_zmarch3d(CaseParams P0, BatchMap M, SweepSel e0, SweepSel e1,
                                                        SweepSel e2, int launch) {
    constexpr int HX = TX + 4, HY = TY + 4, PL = HX * HY;
    constexpr int NPL = (PL + kBlock - 1) / kBlock;
    constexpr int RW = TX / 2 + 1, NR = (TY + 2) * RW;       // red cells of tile + ring per plane
    constexpr int BW = TX / 2, NB = TY * BW;                 // black cells of the tile per plane
    constexpr int NRT = (NR + kBlock - 1) / kBlock, NBT = (NB + kBlock - 1) / kBlock;
    constexpr int NA = 7;               // C, E, W, N, S, T, B
    __shared__ float win[NEQ][4][HY][HX];                    // plane z in slot z & 3
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_ZMARCH, blk);
    if (caseDone(P)) return;
    const int nx = P.nx, ny = P.ny, nz = P.nz, N = P.N;
    const int gxc = (nx + TX - 1) / TX, gyc = (ny + TY - 1) / TY;   // column grid of the case
    const int bx = blk.flat % gxc, byz = blk.flat / gxc;
    const int by = byz % gyc, bz = byz / gyc;
    const SweepSel eqs[3] = {e0, e1, e2};
    const float *ein[NEQ], *esrc[NEQ];
    float *eout[NEQ];
    float erlx[NEQ];
#pragma unroll
    for (int q = 0; q < NEQ; ++q) {
        float *fld = slotField(P, eqs[q].slot), *alt = slotAlt(P, eqs[q].slot);
        const bool odd = ((launch + eqs[q].flip) & 1) != 0;
        ein[q] = odd ? alt : fld;
        eout[q] = odd ? fld : alt;
        esrc[q] = slotSource(P, eqs[q].slot);
        erlx[q] = P.relax[eqs[q].q];
    }
    const float *coef = slotCoef(P, e0.slot);    // momentum or temperature planes
    const int zc = P.zcM;                    // planes per chunk (at most ZC)
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * zc;
    const int z1 = min(z0 + zc, nz);
    const int tid = threadIdx.x;

    auto wslot = [](int z) { return z & 3; };
    auto loadPlane = [&](int z, float (&r)[NEQ][NPL]) {
#pragma unroll
        for (int m = 0; m < NPL; ++m) {
            const int q = tid + m * kBlock;
            const int ly = q / HX, lx = q - ly * HX;
            const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
            const bool in = q < PL && z >= 0 && z < nz && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
            const int g = gx + nx * (gy + ny * z);
#pragma unroll
            for (int e = 0; e < NEQ; ++e) r[e][m] = in ? ein[e][g] : 0.0f;
        }
    };
    auto storePlane = [&](int z, const float (&r)[NEQ][NPL]) {
#pragma unroll
        for (int e = 0; e < NEQ; ++e) {
            float *w = &win[e][wslot(z)][0][0];
#pragma unroll
            for (int m = 0; m < NPL; ++m) {
                const int q = tid + m * kBlock;
                if (q < PL) w[q] = r[e][m];
            }
        }
    };

    // the operands of one relaxed cell: its coefficients and the equations' sources
    struct Ops {
        float a[NA];
        float b[NEQ];
    };
    // cell m of this thread in the red (ring) or black (tile) set of plane z; false if off the grid
    auto redCell = [&](int z, int m, int &gx, int &gy) {
        const int q = tid + m * kBlock;
        const int r = q / RW, kk = q - r * RW;
        gy = y0 - 1 + r;
        gx = x0 - 1 + 2 * kk + ((x0 - 1 + gy + z) & 1);
        return q < NR && z >= 0 && z < nz && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    };
    auto blackCell = [&](int z, int m, int &gx, int &gy) {
        const int q = tid + m * kBlock;
        const int r = q / BW, kk = q - r * BW;
        gy = y0 + r;
        gx = x0 + 2 * kk + ((x0 + gy + z + 1) & 1);
        return q < NB && z < nz && gx < nx && gy < ny;
    };
    auto loadOps = [&](bool ok, int gx, int gy, int z, Ops &o) {
        const int c = gx + nx * (gy + ny * z);
#pragma unroll
        for (int s = 0; s < 7; ++s) o.a[s] = ok ? coef[c + N * s] : 0.0f;
#pragma unroll
        for (int e = 0; e < NEQ; ++e) o.b[e] = ok ? esrc[e][c] : 0.0f;
    };
    auto loadRed = [&](int z, Ops (&o)[NRT]) {
#pragma unroll
        for (int m = 0; m < NRT; ++m) {
            int gx, gy;
            const bool ok = redCell(z, m, gx, gy);
            loadOps(ok, gx, gy, z, o[m]);
        }
    };
    auto loadBlack = [&](int z, Ops (&o)[NBT]) {
#pragma unroll
        for (int m = 0; m < NBT; ++m) {
            int gx, gy;
            const bool ok = blackCell(z, m, gx, gy);
            loadOps(ok, gx, gy, z, o[m]);
        }
    };

    // issue the loads of the first four planes and the first operands, then decide which
    // equations sweep; plane z0+2 stays in flight, in registers, until the first step stores it
    float pend[NEQ][NPL];
    {
        float reg[4][NEQ][NPL];
#pragma unroll
        for (int s = 0; s < 4; ++s) loadPlane(z0 - 2 + s, reg[s]);
        loadPlane(z0 + 2, pend);
#pragma unroll
        for (int s = 0; s < 4; ++s) storePlane(z0 - 2 + s, reg[s]);
    }
    bool run[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; ++e) run[e] = launchRuns(P, blk, eqs[e].slot, launch, eqs[e].kPrev, blk.n, P.tol[eqs[e].q]);
    __syncthreads();

    double acc[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; ++e) acc[e] = 0.0;
    auto relax = [&](int gx, int gy, int z, bool owned, const Ops &o) {
        const int lx = gx - x0 + 2, ly = gy - y0 + 2;
        const float aC = o.a[0], aE = o.a[1], aW = o.a[2], aN = o.a[3], aS = o.a[4], aT = o.a[5], aB = o.a[6];
#pragma unroll
        for (int e = 0; e < NEQ; ++e) {
            if (!run[e]) continue;
            float(*wc)[HX] = win[e][wslot(z)];
            const float phi_C = wc[ly][lx];
            const float phi_E = wc[ly][lx + 1], phi_W = wc[ly][lx - 1];
            const float phi_N = wc[ly + 1][lx], phi_S = wc[ly - 1][lx];
            const float phi_T = (z != nz - 1) ? win[e][wslot(z + 1)][ly][lx] : 0.0f;
            const float phi_B = (z != 0) ? win[e][wslot(z - 1)][ly][lx] : 0.0f;
            float newPhi = o.b[e] - aE * phi_E - aW * phi_W - aN * phi_N - aS * phi_S;
            newPhi -= aT * phi_T + aB * phi_B;
            newPhi /= aC;
            const float d = erlx[e] * (newPhi - phi_C);
            wc[ly][lx] = phi_C + d;
            if (owned) acc[e] = sqAdd(acc[e], d);
        }
    };
    auto redPlane = [&](int z, const Ops (&o)[NRT]) {
        const bool ownPlane = z >= z0 && z < z1;
#pragma unroll
        for (int m = 0; m < NRT; ++m) {
            int gx, gy;
            if (redCell(z, m, gx, gy))
                relax(gx, gy, z, ownPlane && gx >= x0 && gx < x0 + TX && gy >= y0 && gy < y0 + TY, o[m]);
        }
    };
    auto blackPlane = [&](int z, const Ops (&o)[NBT]) {
#pragma unroll
        for (int m = 0; m < NBT; ++m) {
            int gx, gy;
            if (blackCell(z, m, gx, gy)) relax(gx, gy, z, true, o[m]);
        }
    };

    // The red and black operands of one plane are loaded in the same step, so the second load
    // finds the cache lines the first brought in (the cells of a colour sit two floats apart);
    // the black ones are used two steps later.
    Ops red[NRT], blk0[NBT], blk1[NBT];
    {
        Ops r0[NRT];
        loadRed(z0 - 1, r0);
        loadRed(z0, red);
        loadBlack(z0, blk0);
        redPlane(z0 - 1, r0);
        redPlane(z0, red);
    }
    loadRed(z0 + 1, red);
    loadBlack(z0 + 1, blk1);
    // write plane z of the tile
    auto writePlane = [&](int z) {
        for (int q = tid; q < TX * TY; q += kBlock) {
            const int r = q / TX, xx = q - r * TX;
            const int gx = x0 + xx, gy = y0 + r;
            if (gx < nx && gy < ny) {
#pragma unroll
                for (int e = 0; e < NEQ; ++e) eout[e][gx + nx * (gy + ny * z)] = win[e][wslot(z)][r + 2][xx + 2];
            }
        }
    };
    __syncthreads();
    // Step t: store old plane t+2 over plane t-2 (read by no later phase), barrier, red(t+1)
    // (planes t..t+2), barrier, black(t) (planes t-1..t+1), barrier, write plane t.  The store of
    // step t+1 (over plane t-1) follows black(t) by a barrier.
    for (int t = z0; t < z1; ++t) {
        // old plane t+2 (needed by red(t+1)) was loaded during the previous step; plane t+3 and the
        // operands of the next step load during this one
        storePlane(t + 2, pend);
        loadPlane(t + 3, pend);
        Ops redN[NRT], blkN[NBT];
        loadRed(t + 2, redN);
        loadBlack(t + 2, blkN);
        __syncthreads();
        redPlane(t + 1, red);
        __syncthreads();
        blackPlane(t, blk0);
        __syncthreads();
        writePlane(t);
#pragma unroll
        for (int m = 0; m < NRT; ++m) red[m] = redN[m];
#pragma unroll
        for (int m = 0; m < NBT; ++m) {
            blk0[m] = blk1[m];
            blk1[m] = blkN[m];
        }
    }
    {
        double tot[NEQ];
        blockSum<NEQ>(acc, tot);
#pragma unroll
        for (int e = 0; e < NEQ; ++e)
            if (tid == e && run[e]) sweepPart(P, eqs[e].slot, launch, 0)[blk.flat] = tot[e];
    }
}

// One red-black sweep of the p' system per launch, on TX x TY columns of ZC planes.  As in
// k_rb_zmarch3d the column marches up with a rolling window of the field (four planes, 2-cell halo),
// relaxing at step t the red cells of plane t+1 (tile plus a 1-cell ring) and then the black cells of
// plane t; here the matrix also comes through LDS: the stored links towards +x and +y (three-plane
// windows) and towards +z (four planes) are staged over the same region as the field with coalesced
// loads, one step ahead, so each value is fetched once per plane instead of once per cell that reads
// it (a cell's -x, -y and -z links are the +x, +y and +z links of its west, south and lower
// neighbours, bit for bit, kernels.cuh:517-529).  Only the sources are read per cell, into registers.
template <int TX, int TY, int ZC, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_rb_zmarch3d_p(CaseParams P0, BatchMap M, SweepSel e, int launch) {
    constexpr int HX = TX + 4, HY = TY + 4, PL = HX * HY;
    constexpr int NPL = (PL + kBlock - 1) / kBlock;
    constexpr int RW = TX / 2 + 1, NR = (TY + 2) * RW;       // red cells of tile + ring per plane
    constexpr int BW = TX / 2, NB = TY * BW;                 // black cells of the tile per plane
    constexpr int NRT = (NR + kBlock - 1) / kBlock, NBT = (NB + kBlock - 1) / kBlock;
    __shared__ float win[4][HY][HX];                          // p', plane z in slot z & 3
    __shared__ float lT[4][HY][HX];                           // links towards +z, slot z & 3
    __shared__ float lE[3][HY][HX], lN[3][HY][HX];            // towards +x, +y, slot (z + 3) % 3
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_ZMARCH_P, blk);
    if (caseDone(P)) return;
    const int nx = P.nx, ny = P.ny, nz = P.nz, N = P.N;
    const int gxc = (nx + TX - 1) / TX, gyc = (ny + TY - 1) / TY;   // column grid of the case
    const int bx = blk.flat % gxc, byz = blk.flat / gxc;
    const int by = byz % gyc, bz = byz / gyc;
    float *fld = slotField(P, e.slot), *alt = slotAlt(P, e.slot);
    const bool odd = ((launch + e.flip) & 1) != 0;
    const float *in = odd ? alt : fld;
    float *out = odd ? fld : alt;
    const float rlx = P.relax[e.q];
    const int zc = P.zcP;                    // planes per chunk (at most ZC)
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * zc;
    const int z1 = min(z0 + zc, nz);
    const int tid = threadIdx.x;
    // raw buffer loads: an offset past the array reads 0, so no branch surrounds a load
    const int nbytes = N * 4;
    const __amdgpu_buffer_rsrc_t bIn = bufferOf(in, nbytes), bSrc = bufferOf(slotSource(P, e.slot), nbytes);
    const __amdgpu_buffer_rsrc_t bE = bufferOf(P.pcoef + (size_t)N * sE, nbytes);
    const __amdgpu_buffer_rsrc_t bN = bufferOf(P.pcoef + (size_t)N * sN, nbytes);
    const __amdgpu_buffer_rsrc_t bT = bufferOf(P.pcoef + (size_t)N * sT, nbytes);
    auto s3 = [](int z) { return (z + 3) % 3; };              // z >= -3

    // the staged region of plane z (0 off the grid and outside [0, nz))
    auto loadPlane = [&](__amdgpu_buffer_rsrc_t b, int z, float (&r)[NPL]) {
#pragma unroll
        for (int m = 0; m < NPL; ++m) {
            const int q = tid + m * kBlock;
            const int ly = q / HX, lx = q - ly * HX;
            const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
            const bool ok = (q < PL) & ((unsigned)z < (unsigned)nz) & ((unsigned)gx < (unsigned)nx)
                            & ((unsigned)gy < (unsigned)ny);
            r[m] = bufferLoad(b, ok ? 4 * (gx + nx * (gy + ny * z)) : kNoOffset);
        }
    };
    auto storePlane = [&](float *w, const float (&r)[NPL]) {
#pragma unroll
        for (int m = 0; m < NPL; ++m) {
            const int q = tid + m * kBlock;
            if (q < PL) w[q] = r[m];
        }
    };
    struct Links {
        float e[NPL], n[NPL], t[NPL];
    };
    auto loadLinks = [&](int z, Links &l) {
        loadPlane(bE, z, l.e);
        loadPlane(bN, z, l.n);
        loadPlane(bT, z, l.t);
    };
    auto storeLinks = [&](int z, const Links &l) {
        storePlane(&lE[s3(z)][0][0], l.e);
        storePlane(&lN[s3(z)][0][0], l.n);
        storePlane(&lT[z & 3][0][0], l.t);
    };
    // cell m of this thread in the red (ring) or black (tile) set of plane z; false if off the grid
    auto redCell = [&](int z, int m, int &gx, int &gy) {
        const int q = tid + m * kBlock;
        const int r = q / RW, kk = q - r * RW;
        gy = y0 - 1 + r;
        gx = x0 - 1 + 2 * kk + ((x0 - 1 + gy + z) & 1);
        return q < NR && z >= 0 && z < nz && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    };
    auto blackCell = [&](int z, int m, int &gx, int &gy) {
        const int q = tid + m * kBlock;
        const int r = q / BW, kk = q - r * BW;
        gy = y0 + r;
        gx = x0 + 2 * kk + ((x0 + gy + z + 1) & 1);
        return q < NB && z < nz && gx < nx && gy < ny;
    };
    auto loadSrcRed = [&](int z, float (&b)[NRT]) {
#pragma unroll
        for (int m = 0; m < NRT; ++m) {
            int gx, gy;
            const bool ok = redCell(z, m, gx, gy);
            b[m] = bufferLoad(bSrc, ok ? 4 * (gx + nx * (gy + ny * z)) : kNoOffset);
        }
    };
    auto loadSrcBlack = [&](int z, float (&b)[NBT]) {
#pragma unroll
        for (int m = 0; m < NBT; ++m) {
            int gx, gy;
            const bool ok = blackCell(z, m, gx, gy);
            b[m] = bufferLoad(bSrc, ok ? 4 * (gx + nx * (gy + ny * z)) : kNoOffset);
        }
    };
    auto writePlane = [&](int z, const float *from) {
        for (int q = tid; q < TX * TY; q += kBlock) {
            const int r = q / TX, xx = q - r * TX;
            const int gx = x0 + xx, gy = y0 + r;
            if (gx < nx && gy < ny) {
                const int g = gx + nx * (gy + ny * z);
                out[g] = from ? from[g] : win[z & 3][r + 2][xx + 2];
            }
        }
    };

    double acc = 0.0;
    auto relax = [&](int gx, int gy, int z, bool owned, float b) {
        const int lx = gx - x0 + 2, ly = gy - y0 + 2;
        const int c3 = s3(z);
        const float aE = lE[c3][ly][lx], aW = lE[c3][ly][lx - 1];
        const float aN = lN[c3][ly][lx], aS = lN[c3][ly - 1][lx];
        const float aT = lT[z & 3][ly][lx], aB = lT[(z - 1) & 3][ly][lx];
        float aC = -(aE + aW + aN + aS);
        aC += -(aT + aB);
        float(*wc)[HX] = win[z & 3];
        const float phi_C = wc[ly][lx];
        const float phi_E = wc[ly][lx + 1], phi_W = wc[ly][lx - 1];
        const float phi_N = wc[ly + 1][lx], phi_S = wc[ly - 1][lx];
        const float phi_T = (z != nz - 1) ? win[(z + 1) & 3][ly][lx] : 0.0f;
        const float phi_B = (z != 0) ? win[(z - 1) & 3][ly][lx] : 0.0f;
        float newPhi = b - aE * phi_E - aW * phi_W - aN * phi_N - aS * phi_S;
        newPhi -= aT * phi_T + aB * phi_B;
        newPhi /= aC;
        const float d = rlx * (newPhi - phi_C);
        wc[ly][lx] = phi_C + d;
        if (owned) acc = sqAdd(acc, d);
    };
    auto redPlane = [&](int z, const float (&b)[NRT]) {
        const bool ownPlane = z >= z0 && z < z1;
#pragma unroll
        for (int m = 0; m < NRT; ++m) {
            int gx, gy;
            if (redCell(z, m, gx, gy))
                relax(gx, gy, z, ownPlane && gx >= x0 && gx < x0 + TX && gy >= y0 && gy < y0 + TY, b[m]);
        }
    };
    auto blackPlane = [&](int z, const float (&b)[NBT]) {
#pragma unroll
        for (int m = 0; m < NBT; ++m) {
            int gx, gy;
            if (blackCell(z, m, gx, gy)) relax(gx, gy, z, true, b[m]);
        }
    };

    // field planes z0-2..z0+1 and the links of planes z0-2 (+z only), z0-1 and z0 now; the field
    // of plane z0+2 and the links of plane z0+1 stay in flight until the first step stores them
    float pend[NPL];
    Links pendL;
    float redB[NRT], blkB[NBT], redN[NRT], blkN[NBT];
    {
        float f[4][NPL];
        Links l[2];
        float tlow[NPL];
#pragma unroll
        for (int s = 0; s < 4; ++s) loadPlane(bIn, z0 - 2 + s, f[s]);
        loadPlane(bT, z0 - 2, tlow);
        loadLinks(z0 - 1, l[0]);
        loadLinks(z0, l[1]);
        loadPlane(bIn, z0 + 2, pend);
        loadLinks(z0 + 1, pendL);
        float r0[NRT];
        loadSrcRed(z0 - 1, r0);
        loadSrcRed(z0, redB);
        loadSrcBlack(z0, blkB);
        // the stop decision, with these loads in flight
        if (!launchRuns(P, blk, e.slot, launch, e.kPrev, blk.n, P.tol[e.q])) {
            // a stopped call: the input unchanged (keeps the ping-pong parity)
            for (int z = z0; z < z1; ++z) writePlane(z, in);
            return;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) storePlane(&win[(z0 - 2 + s) & 3][0][0], f[s]);
        storePlane(&lT[(z0 - 2) & 3][0][0], tlow);
        storeLinks(z0 - 1, l[0]);
        storeLinks(z0, l[1]);
        __syncthreads();
        redPlane(z0 - 1, r0);
    }
    redPlane(z0, redB);
    loadSrcRed(z0 + 1, redB);
    __syncthreads();
    // Step t: store old plane t+2 (over t-2) and the links of plane t+1 (over those of t-2, and of
    // t-3 for +z), barrier, red(t+1) (field planes t..t+2, links of t+1 and +z of t), barrier,
    // black(t) (field t-1..t+1, links of t and +z of t-1), barrier, write plane t.
    for (int t = z0; t < z1; ++t) {
        storePlane(&win[(t + 2) & 3][0][0], pend);
        storeLinks(t + 1, pendL);
        loadPlane(bIn, t + 3, pend);
        loadLinks(t + 2, pendL);
        loadSrcRed(t + 2, redN);
        loadSrcBlack(t + 1, blkN);
        __syncthreads();
        redPlane(t + 1, redB);
        __syncthreads();
        blackPlane(t, blkB);
        __syncthreads();
        writePlane(t, nullptr);
#pragma unroll
        for (int m = 0; m < NRT; ++m) redB[m] = redN[m];
#pragma unroll
        for (int m = 0; m < NBT; ++m) blkB[m] = blkN[m];
    }
    {
        double v[1] = {acc}, tot[1];
        blockSum<1>(v, tot);
        if (tid == 0) sweepPart(P, e.slot, launch, 0)[blk.flat] = tot[0];
    }
}

// Two red-black sweeps of the p' system per launch, on the same TX x TY x ZC columns as
// k_rb_zmarch3d_p (so the workgroup counts and partial records are the same).  The field and the
// +x, +y, +z links are staged over the tile plus a 4-cell halo; at step t the column runs four
// phases, each behind a barrier:
//   red1(t+1) on ring <= 3   reads the old black values of planes t..t+2
//   black1(t) on ring <= 2   reads red1 of planes t-1..t+1
//   red2(t-1) on ring <= 1   reads black1 of planes t-2..t
//   black2(t-2) on the tile  reads red2 of planes t-3..t-1, then plane t-2 is written
// (ring r: the cells within r of the tile in x and y).  Each phase's ring is the set the next one
// reads, and below and above the column the phases start 3..0 planes early and end 3..0 planes
// late (red1 on [z0-3, z1+2], black1 on [z0-2, z1+1], red2 on [z0-1, z1], black2 on [z0, z1)), so
// every value is recomputed from the same inputs its owning column uses and the launch equals two
// of the reference's red-pass-then-black-pass sweeps bit for bit.  Live planes at step t: the
// field t-3..t+2 (six slots), +z links t-3..t+1 (five), +x/+y links t-2..t+1 (four).
template <int TX, int TY, int ZC, bool BATCH>
__global__ __launch_bounds__(kBlock) void k_rb_zmarch3d_p2(CaseParams P0, BatchMap M, SweepSel e, int launch) {
    constexpr int G = 4;                                      // staged halo
    constexpr int HX = TX + 2 * G, HY = TY + 2 * G, PL = HX * HY;
    constexpr int NPL = (PL + kBlock - 1) / kBlock;
    constexpr int NC = ((TX + 6) / 2 * (TY + 6) + kBlock - 1) / kBlock;   // cells per thread, largest ring
    constexpr int FS = 6, TS = 5, ES = 4;
    static_assert(TX % 2 == 0, "colour rows need an even tile width");
    __shared__ float win[FS][HY][HX];                         // p', plane z in slot fs(z)
    __shared__ float lT[TS][HY][HX];                          // links towards +z, slot ts(z)
    __shared__ float lE[ES][HY][HX], lN[ES][HY][HX];          // towards +x, +y, slot z & 3
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_ZMARCH_P, blk);
    if (caseDone(P)) return;
    const int nx = P.nx, ny = P.ny, nz = P.nz, N = P.N;
    const int gxc = (nx + TX - 1) / TX, gyc = (ny + TY - 1) / TY;
    const int bx = blk.flat % gxc, byz = blk.flat / gxc;
    const int by = byz % gyc, bz = byz / gyc;
    float *fld = slotField(P, e.slot), *alt = slotAlt(P, e.slot);
    const bool odd = ((launch + e.flip) & 1) != 0;
    const float *in = odd ? alt : fld;
    float *out = odd ? fld : alt;
    const float rlx = P.relax[e.q];
    const int zc = P.zcP;                    // planes per chunk (at most ZC)
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * zc;
    const int z1 = min(z0 + zc, nz);
    const int tid = threadIdx.x;
    const int nbytes = N * 4;
    const __amdgpu_buffer_rsrc_t bIn = bufferOf(in, nbytes), bSrc = bufferOf(slotSource(P, e.slot), nbytes);
    const __amdgpu_buffer_rsrc_t bE = bufferOf(P.pcoef + (size_t)N * sE, nbytes);
    const __amdgpu_buffer_rsrc_t bN = bufferOf(P.pcoef + (size_t)N * sN, nbytes);
    const __amdgpu_buffer_rsrc_t bT = bufferOf(P.pcoef + (size_t)N * sT, nbytes);
    auto fs = [](int z) { return (z + 12) % FS; };            // z >= -12
    auto ts = [](int z) { return (z + 10) % TS; };            // z >= -10

    auto loadPlane = [&](__amdgpu_buffer_rsrc_t b, int z, float (&r)[NPL]) {
#pragma unroll
        for (int m = 0; m < NPL; ++m) {
            const int q = tid + m * kBlock;
            const int ly = q / HX, lx = q - ly * HX;
            const int gx = x0 - G + lx, gy = y0 - G + ly;
            const bool ok = (q < PL) & ((unsigned)z < (unsigned)nz) & ((unsigned)gx < (unsigned)nx)
                            & ((unsigned)gy < (unsigned)ny);
            r[m] = bufferLoad(b, ok ? 4 * (gx + nx * (gy + ny * z)) : kNoOffset);
        }
    };
    auto storePlane = [&](float *w, const float (&r)[NPL]) {
#pragma unroll
        for (int m = 0; m < NPL; ++m) {
            const int q = tid + m * kBlock;
            if (q < PL) w[q] = r[m];
        }
    };
    struct Links {
        float e[NPL], n[NPL], t[NPL];
    };
    auto loadLinks = [&](int z, Links &l) {
        loadPlane(bE, z, l.e);
        loadPlane(bN, z, l.n);
        loadPlane(bT, z, l.t);
    };
    auto storeLinks = [&](int z, const Links &l) {
        storePlane(&lE[z & 3][0][0], l.e);
        storePlane(&lN[z & 3][0][0], l.n);
        storePlane(&lT[ts(z)][0][0], l.t);
    };
    // cell q of colour c (0 red, 1 black: (x + y + z) even / odd) within ring r of the tile in
    // plane z, row-major over the ring's rectangle; false past the set or off the grid
    auto cellAt = [&](int r, int c, int z, int q, int &gx, int &gy) {
        const int w = TX / 2 + r;
        const int row = q / w, kk = q - row * w;
        gy = y0 - r + row;
        const int xs = x0 - r;
        gx = xs + 2 * kk + ((xs + gy + z + c) & 1);
        return q < (TY + 2 * r) * w && z >= 0 && z < nz && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    };
    auto loadSrc = [&](int r, int c, int z, bool on, float (&b)[NC]) {
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            int gx, gy;
            const bool ok = on && cellAt(r, c, z, tid + m * kBlock, gx, gy);
            b[m] = bufferLoad(bSrc, ok ? 4 * (gx + nx * (gy + ny * z)) : kNoOffset);
        }
    };
    auto writePlane = [&](int z, const float *from) {
        for (int q = tid; q < TX * TY; q += kBlock) {
            const int r = q / TX, xx = q - r * TX;
            const int gx = x0 + xx, gy = y0 + r;
            if (gx < nx && gy < ny) {
                const int g = gx + nx * (gy + ny * z);
                out[g] = from ? from[g] : win[fs(z)][r + G][xx + G];
            }
        }
    };

    double acc0 = 0.0, acc1 = 0.0;
    // the arithmetic of k_rb_zmarch3d_p's relax, term for term
    auto relax = [&](int gx, int gy, int z, bool owned, float b, bool second) {
        const int lx = gx - x0 + G, ly = gy - y0 + G;
        const int c4 = z & 3;
        const float aE = lE[c4][ly][lx], aW = lE[c4][ly][lx - 1];
        const float aN = lN[c4][ly][lx], aS = lN[c4][ly - 1][lx];
        const float aT = lT[ts(z)][ly][lx], aB = lT[ts(z - 1)][ly][lx];
        float aC = -(aE + aW + aN + aS);
        aC += -(aT + aB);
        float(*wc)[HX] = win[fs(z)];
        const float phi_C = wc[ly][lx];
        const float phi_E = wc[ly][lx + 1], phi_W = wc[ly][lx - 1];
        const float phi_N = wc[ly + 1][lx], phi_S = wc[ly - 1][lx];
        const float phi_T = (z != nz - 1) ? win[fs(z + 1)][ly][lx] : 0.0f;
        const float phi_B = (z != 0) ? win[fs(z - 1)][ly][lx] : 0.0f;
        float newPhi = b - aE * phi_E - aW * phi_W - aN * phi_N - aS * phi_S;
        newPhi -= aT * phi_T + aB * phi_B;
        newPhi /= aC;
        const float d = rlx * (newPhi - phi_C);
        wc[ly][lx] = phi_C + d;
        if (owned) {
            if (second) acc1 = sqAdd(acc1, d);
            else acc0 = sqAdd(acc0, d);
        }
    };
    auto phase = [&](int r, int c, int z, bool on, const float (&b)[NC], bool second) {
        if (!on) return;
        const bool ownPlane = z >= z0 && z < z1;
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            int gx, gy;
            if (cellAt(r, c, z, tid + m * kBlock, gx, gy))
                relax(gx, gy, z, ownPlane && gx >= x0 && gx < x0 + TX && gy >= y0 && gy < y0 + TY, b[m], second);
        }
    };

    // steps t = tA .. z1+1 (none is needed below plane -1); a phase runs when its plane is in its
    // range: red1 always, black1 from z0-2, red2 from z0-1, black2 from z0
    const int tA = max(z0 - 4, -1);
    float pend[NPL];
    Links pendL;
    float sR1[NC], sB1[NC], sR2[NC], sB2[NC];
    {
        float f0[NPL], f1[NPL], tl[NPL];
        loadPlane(bIn, tA, f0);
        loadPlane(bIn, tA + 1, f1);
        loadPlane(bT, tA, tl);
        loadPlane(bIn, tA + 2, pend);
        loadLinks(tA + 1, pendL);
        loadSrc(3, 0, tA + 1, true, sR1);
        loadSrc(2, 1, tA, tA >= z0 - 2, sB1);
        loadSrc(1, 0, tA - 1, tA >= z0, sR2);
        loadSrc(0, 1, tA - 2, tA >= z0 + 2, sB2);
        // the stop decision, with these loads in flight
        if (!launchRuns(P, blk, e.slot, launch, e.kPrev, blk.n, P.tol[e.q])) {
            for (int z = z0; z < z1; ++z) writePlane(z, in);      // keeps the ping-pong parity
            return;
        }
        storePlane(&win[fs(tA)][0][0], f0);
        storePlane(&win[fs(tA + 1)][0][0], f1);
        storePlane(&lT[ts(tA)][0][0], tl);
    }
    for (int t = tA; t <= z1 + 1; ++t) {
        storePlane(&win[fs(t + 2)][0][0], pend);
        storeLinks(t + 1, pendL);
        loadPlane(bIn, t + 3, pend);
        loadLinks(t + 2, pendL);
        float nR1[NC], nB1[NC], nR2[NC], nB2[NC];
        loadSrc(3, 0, t + 2, true, nR1);
        loadSrc(2, 1, t + 1, t + 1 >= z0 - 2, nB1);
        loadSrc(1, 0, t, t + 1 >= z0, nR2);
        loadSrc(0, 1, t - 1, t + 1 >= z0 + 2, nB2);
        __syncthreads();
        phase(3, 0, t + 1, true, sR1, false);
        __syncthreads();
        phase(2, 1, t, t >= z0 - 2, sB1, false);
        __syncthreads();
        phase(1, 0, t - 1, t >= z0, sR2, true);
        __syncthreads();
        phase(0, 1, t - 2, t >= z0 + 2, sB2, true);
        __syncthreads();
        if (t - 2 >= z0) writePlane(t - 2, nullptr);
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            sR1[m] = nR1[m];
            sB1[m] = nB1[m];
            sR2[m] = nR2[m];
            sB2[m] = nB2[m];
        }
    }
    {
        double v[2] = {acc0, acc1}, tot[2];
        blockSum<2>(v, tot);
        if (tid == 0) {
            sweepPart(P, e.slot, launch, 0)[blk.flat] = tot[0];
            sweepPart(P, e.slot, launch, 1)[blk.flat] = tot[1];
        }
    }
}

// z planes per column chunk: 32 (Z_ZC, ZP_ZC), halved while a launch would have fewer than 512
// workgroups (two per CU), down to 4.  A column's march is a chain of dependent plane steps, so on
// grids with few columns shorter chunks spread it over more CUs; the planes each chunk recomputes
// at its edges cost less than the idle CUs (64^3: 32 columns of 32 planes took 40.6 us per p'
// sweep).  256^3 keeps 32.
int zmarch_chunk(const CaseParams &P, bool pc) {
    const long long cols = pc ? (long long)((P.nx + ZP_TX - 1) / ZP_TX) * ((P.ny + ZP_TY - 1) / ZP_TY)
                              : (long long)((P.nx + Z_TX - 1) / Z_TX) * ((P.ny + Z_TY - 1) / Z_TY);
    int zc = pc ? ZP_ZC : Z_ZC;
    while (zc > 4 && cols * ((P.nz + zc - 1) / zc) < 512) zc /= 2;
    return zc;
}

int zmarch_blocks(const CaseParams &P, bool pc) {
    const int zc = pc ? P.zcP : P.zcM;
    if (pc) return ((P.nx + ZP_TX - 1) / ZP_TX) * ((P.ny + ZP_TY - 1) / ZP_TY) * ((P.nz + zc - 1) / zc);
    return ((P.nx + Z_TX - 1) / Z_TX) * ((P.ny + Z_TY - 1) / Z_TY) * ((P.nz + zc - 1) / zc);
}

int zmarch_dimx(const CaseParams &P, bool pc) { return pc ? (P.nx + ZP_TX - 1) / ZP_TX : (P.nx + Z_TX - 1) / Z_TX; }

template <int NEQ, bool PC, bool BATCH>
static void launchZ(const Lc &L, const SweepSel *e, int launch) {
    const CaseParams &P = *L.P;
    const dim3 grid = BATCH ? dim3((unsigned)L.total[PC ? MAP_ZMARCH_P : MAP_ZMARCH]) : dim3((unsigned)zmarch_blocks(P, PC));
    if constexpr (PC) {
        if (e[0].k >= 2) k_rb_zmarch3d_p2<ZP_TX, ZP_TY, ZP_ZC, BATCH><<<grid, kBlock, 0, L.s>>>(P, L.M, e[0], launch);
        else k_rb_zmarch3d_p<ZP_TX, ZP_TY, ZP_ZC, BATCH><<<grid, kBlock, 0, L.s>>>(P, L.M, e[0], launch);
    } else
        k_rb_zmarch3d<Z_TX, Z_TY, Z_ZC, NEQ, BATCH><<<grid, kBlock, 0, L.s>>>(P, L.M, e[0], e[NEQ > 1 ? 1 : 0],
                                                                              e[NEQ > 2 ? 2 : 0], launch);
}

template <bool BATCH>
static void launchZ(const Lc &L, int kind, const SweepSel *e, int neq, int launch) {
    if (kind == COEF_PCORR) launchZ<1, true, BATCH>(L, e, launch);
    else if (neq == 3) launchZ<3, false, BATCH>(L, e, launch);
    else if (neq == 2) launchZ<2, false, BATCH>(L, e, launch);
    else launchZ<1, false, BATCH>(L, e, launch);
}

void launch_rb_sweep3d(const Lc &L, int kind, const SweepSel *e, int neq, int launch) {
    if (L.batch) launchZ<true>(L, kind, e, neq, launch);
    else launchZ<false>(L, kind, e, neq, launch);
}

}  // namespace ifl
