// mg3d.hip - gfx950 kernels of the multigrid p' solve on 3D grids (IFL_PSOLVE_MG).
//
// Level-wide kernels, one thread per cell (or coarse cell):
//   k_mg3_coarsen  the matrix of level l+1 from level l: the sums of the fine couplings across each
//                  coarse face, the coarse outlet term and the centre; a coarse cell also sums its
//                  west, south and bottom neighbours' +x/+y/+z couplings (the same sums in the same
//                  order, so the same bits as the neighbour's own)
//   k_mg3_sweep    one colour of an unrelaxed red-black sweep, in place (a cell reads only the other
//                  colour, so the result does not depend on the order of the cells)
//   k_mg3_restrict the FP64 residual of the 2x2x2 children summed in FP64 (order k, j, i) and rounded
//                  once into the coarse source; the coarse correction starts at 0
//   k_mg3_prolong  x += over * x_parent
// Level 0 is the case's p' system in place: x = p' (zeroed by k_pcorr_assemble), b = b', the +x/+y/+z
// couplings and the centre of the assembled p' matrix (kernels.cuh:502-540) plus, next to outlet
// faces, the coupling rho A^2 / aC to the ghost p' = 0 of the velocity correction (kernels.cuh:713-727).
// The arithmetic is the oracle's (simple_oracle.hpp, mgBuild .. mgCycle) expression by expression
// (-ffp-contract=off).
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_mg.hpp"

namespace ifl {

namespace {

constexpr int kMg3Block = 256;

// Outlet term of level-0 cell (i, j, k): faces in the order E, W, N, S, T, B.
__device__ __forceinline__ float mg3Outlet0(const CaseParams &P, int i, int j, int k, int c) {
    const bool at[6] = {i == P.nx - 1, i == 0, j == P.ny - 1, j == 0, k == P.nz - 1, k == 0};
    bool any = false;
#pragma unroll
    for (int f = 0; f < 6; ++f) any |= at[f] && P.bct[f] == IFL_OUTLET;
    float dg = 0.0f;
    if (any) {
        const float rinv = P.rinv[c];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            if (!at[f] || P.bct[f] != IFL_OUTLET) continue;
            const float A = f < 2 ? P.Ax : f < 4 ? P.Ay : P.Az;
            dg += P.rho * A * A * rinv;
        }
    }
    return dg;
}

// The planes of a level (level 0: the case's p' system).
struct Mg3View {
    int nx, ny, nz;
    float *x;
    const float *b, *E, *N, *T, *C, *dg;
};

template <bool L0>
__device__ __forceinline__ Mg3View viewOf(const CaseParams &P, const MgLevel3 &F) {
    Mg3View v;
    v.nx = F.nx;
    v.ny = F.ny;
    v.nz = F.nz;
    if constexpr (L0) {
        const size_t n = (size_t)P.N;
        v.x = P.pc;
        v.b = P.bp;
        v.C = P.pcoef + n * sC;
        v.E = P.pcoef + n * sE;
        v.N = P.pcoef + n * sN;
        v.T = P.pcoef + n * sT;
        v.dg = nullptr;
    } else {
        v.x = F.x;
        v.b = F.b;
        v.C = F.C;
        v.E = F.E;
        v.N = F.N;
        v.T = F.T;
        v.dg = F.dg;
    }
    return v;
}

// Centre of cell c (level 0: the assembled centre plus the outlet term, as mgBuild stores it).
template <bool L0>
__device__ __forceinline__ float centreOf(const CaseParams &P, const Mg3View &v, int i, int j, int k, int c) {
    if constexpr (L0) return v.C[c] + mg3Outlet0(P, i, j, k, c);
    else return v.C[c];
}

template <bool L0>
__device__ __forceinline__ float outletOf(const CaseParams &P, const Mg3View &v, int i, int j, int k, int c) {
    if constexpr (L0) return mg3Outlet0(P, i, j, k, c);
    else return v.dg[c];
}

// ---------------- coarse matrices (mgBuild) ----------------
// rec > 0: this launch also records the p' solve's sweep count (the first launch of a solve).
template <bool L0>
__global__ __launch_bounds__(kMg3Block) void k_mg3_coarsen(CaseParams P, MgLevel3 Fl, MgLevel3 Cl, int rec) {
    if (caseDone(P)) return;
    if (rec > 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        P.st->sm[3].sweeps = rec;
        P.st->sm[3].lastSum = 0.0;
        P.st->sweepsIter[3] += rec;
    }
    const int g = blockIdx.x * kMg3Block + threadIdx.x;
    const int ncx = Cl.nx, ncy = Cl.ny, ncz = Cl.nz;
    if (g >= ncx * ncy * ncz) return;
    const Mg3View F = viewOf<L0>(P, Fl);
    const int I = g % ncx, R = g / ncx, J = R % ncy, K = R / ncy;
    auto fid = [&](int i, int j, int k) { return i + F.nx * (j + F.ny * k); };
    auto lo = [](int Q) { return 2 * Q; };
    auto hiX = [&](int Q) { return min(2 * Q + 1, F.nx - 1); };
    auto hiY = [&](int Q) { return min(2 * Q + 1, F.ny - 1); };
    auto hiZ = [&](int Q) { return min(2 * Q + 1, F.nz - 1); };
    // the coupling across the +x face of coarse cell (I, J, K): the fine +x couplings of its last column
    auto sumE = [&](int I_, int J_, int K_) {
        float e = 0.0f;
        const int ie = hiX(I_);
        for (int k = lo(K_); k <= hiZ(K_); ++k)
            for (int j = lo(J_); j <= hiY(J_); ++j) e += F.E[fid(ie, j, k)];
        return e;
    };
    auto sumN = [&](int I_, int J_, int K_) {
        float n = 0.0f;
        const int j1 = hiY(J_);
        for (int k = lo(K_); k <= hiZ(K_); ++k)
            for (int i = lo(I_); i <= hiX(I_); ++i) n += F.N[fid(i, j1, k)];
        return n;
    };
    auto sumT = [&](int I_, int J_, int K_) {
        float t = 0.0f;
        const int k1 = hiZ(K_);
        for (int j = lo(J_); j <= hiY(J_); ++j)
            for (int i = lo(I_); i <= hiX(I_); ++i) t += F.T[fid(i, j, k1)];
        return t;
    };
    float d = 0.0f;
    for (int k = lo(K); k <= hiZ(K); ++k)
        for (int j = lo(J); j <= hiY(J); ++j)
            for (int i = lo(I); i <= hiX(I); ++i) d += outletOf<L0>(P, F, i, j, k, fid(i, j, k));
    const float e = sumE(I, J, K), n = sumN(I, J, K), t = sumT(I, J, K);
    const float w = I > 0 ? sumE(I - 1, J, K) : 0.0f;
    const float s = J > 0 ? sumN(I, J - 1, K) : 0.0f;
    const float bt = K > 0 ? sumT(I, J, K - 1) : 0.0f;
    float cc = -(e + w + n + s);
    cc += -(t + bt);
    Cl.E[g] = e;
    Cl.N[g] = n;
    Cl.T[g] = t;
    Cl.dg[g] = d;
    Cl.C[g] = cc + d;
}

// ---------------- one colour of a red-black sweep (mgSmooth) ----------------
// rec > 0: this launch also records the p' solve's sweep count (the first launch of a solve).
template <bool L0>
__global__ __launch_bounds__(kMg3Block) void k_mg3_sweep(CaseParams P, MgLevel3 Fl, int colour, int rec) {
    if (caseDone(P)) return;
    if (rec > 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        P.st->sm[3].sweeps = rec;
        P.st->sm[3].lastSum = 0.0;
        P.st->sweepsIter[3] += rec;
    }
    const Mg3View F = viewOf<L0>(P, Fl);
    const int nx = F.nx, ny = F.ny, nz = F.nz;
    const int hx = (nx + 1) >> 1;
    const int g = blockIdx.x * kMg3Block + threadIdx.x;
    const int row = g / hx;
    if (row >= ny * nz) return;
    const int j = row % ny, k = row / ny;
    const int i = 2 * (g - row * hx) + ((colour + j + k) & 1);
    if (i >= nx) return;
    const int pl = nx * ny;
    const int c = i + nx * row;
    float *x = F.x;
    const float xE = i < nx - 1 ? x[c + 1] : 0.0f, xW = i > 0 ? x[c - 1] : 0.0f;
    const float xN = j < ny - 1 ? x[c + nx] : 0.0f, xS = j > 0 ? x[c - nx] : 0.0f;
    const float aW = i > 0 ? F.E[c - 1] : 0.0f, aS = j > 0 ? F.N[c - nx] : 0.0f;
    float v = F.b[c] - F.E[c] * xE - aW * xW - F.N[c] * xN - aS * xS;
    const float xT = k < nz - 1 ? x[c + pl] : 0.0f, xB = k > 0 ? x[c - pl] : 0.0f;
    const float aB = k > 0 ? F.T[c - pl] : 0.0f;
    v -= F.T[c] * xT + aB * xB;
    x[c] = v / centreOf<L0>(P, F, i, j, k, c);
}

// FP64 residual of cell (i, j, k), terms in the order C, E, W, N, S, T, B (mgResidual).
template <bool L0>
__device__ __forceinline__ double residual(const CaseParams &P, const Mg3View &F, int i, int j, int k) {
    const int pl = F.nx * F.ny, c = i + F.nx * (j + F.ny * k);
    const float *x = F.x;
    double r = (double)F.b[c];
    r -= (double)centreOf<L0>(P, F, i, j, k, c) * (double)x[c];
    r -= (double)F.E[c] * (double)(i < F.nx - 1 ? x[c + 1] : 0.0f);
    r -= (double)(i > 0 ? F.E[c - 1] : 0.0f) * (double)(i > 0 ? x[c - 1] : 0.0f);
    r -= (double)F.N[c] * (double)(j < F.ny - 1 ? x[c + F.nx] : 0.0f);
    r -= (double)(j > 0 ? F.N[c - F.nx] : 0.0f) * (double)(j > 0 ? x[c - F.nx] : 0.0f);
    r -= (double)F.T[c] * (double)(k < F.nz - 1 ? x[c + pl] : 0.0f);
    r -= (double)(k > 0 ? F.T[c - pl] : 0.0f) * (double)(k > 0 ? x[c - pl] : 0.0f);
    return r;
}

// ---------------- residual restriction (mgRestrict) ----------------
template <bool L0>
__global__ __launch_bounds__(kMg3Block) void k_mg3_restrict(CaseParams P, MgLevel3 Fl, MgLevel3 Cl) {
    if (caseDone(P)) return;
    const int g = blockIdx.x * kMg3Block + threadIdx.x;
    if (g >= Cl.nx * Cl.ny * Cl.nz) return;
    const Mg3View F = viewOf<L0>(P, Fl);
    const int I = g % Cl.nx, R = g / Cl.nx, J = R % Cl.ny, K = R / Cl.ny;
    const int i1 = min(2 * I + 1, F.nx - 1), j1 = min(2 * J + 1, F.ny - 1), k1 = min(2 * K + 1, F.nz - 1);
    double acc = 0.0;
    for (int k = 2 * K; k <= k1; ++k)
        for (int j = 2 * J; j <= j1; ++j)
            for (int i = 2 * I; i <= i1; ++i) acc += residual<L0>(P, F, i, j, k);
    Cl.b[g] = (float)acc;
    Cl.x[g] = 0.0f;
}

// Level 0's restriction with the fine data staged in LDS: a workgroup owns RX x RY coarse cells of one
// coarse plane (a 2RX x 2RY x 2 fine block) and loads, coalesced, the field over the block with a
// 1-cell halo, the +x/+y/+z couplings it needs (including the west, south and lower neighbours' ones
// that serve as the block's -x, -y, -z couplings), the sources and the centres; cells off the grid
// stage as 0, which is what the residual's boundary terms use.  Each thread then sums its coarse
// cell's children in the order k, j, i, as k_mg3_restrict does.
constexpr int RX = 32, RY = 8;
template <int W_, int H_, int D_>
struct Box {
    static constexpr int W = W_, H = H_, D = D_, NQ = (W * H * D + kMg3Block - 1) / kMg3Block;
};

__global__ __launch_bounds__(kMg3Block) void k_mg3_restrict0(CaseParams P, MgLevel3 Fl, MgLevel3 Cl) {
    constexpr int FX = 2 * RX, FY = 2 * RY;
    constexpr int XW = FX + 2, XH = FY + 2;                   // field: halo 1, planes 2K-1 .. 2K+2
    __shared__ float qx[4][XH][XW];
    __shared__ float qE[2][FY][FX + 1];                       // columns 2I0-1 .. 2I0+FX-1
    __shared__ float qN[2][FY + 1][FX];                       // rows 2J0-1 .. 2J0+FY-1
    __shared__ float qT[3][FY][FX];                           // planes 2K-1 .. 2K+1
    __shared__ float qB[2][FY][FX], qC[2][FY][FX];
    if (caseDone(P)) return;
    const int nx = Fl.nx, ny = Fl.ny, nz = Fl.nz;
    const int cbx = (Cl.nx + RX - 1) / RX, cby = (Cl.ny + RY - 1) / RY;
    const int w = xcdSwizzle((int)blockIdx.x, (int)gridDim.x);
    const int bx = w % cbx, r = w / cbx, by = r % cby, K = r / cby;
    const int fx0 = 2 * RX * bx, fy0 = 2 * RY * by, fz0 = 2 * K;
    const size_t n = (size_t)P.N;
    const int nbytes = P.N * 4;
    const __amdgpu_buffer_rsrc_t bX = bufferOf(P.pc, nbytes), bB = bufferOf(P.bp, nbytes);
    const __amdgpu_buffer_rsrc_t bC = bufferOf(P.pcoef + n * sC, nbytes), bE = bufferOf(P.pcoef + n * sE, nbytes);
    const __amdgpu_buffer_rsrc_t bN = bufferOf(P.pcoef + n * sN, nbytes), bT = bufferOf(P.pcoef + n * sT, nbytes);
    const int tid = threadIdx.x;
    // a W x H x D box starting at fine cell (ax, ay, az) of array b (0 off the grid): every load of a
    // box is issued before any store to LDS
    auto load = [&](auto whd, __amdgpu_buffer_rsrc_t b, int ax, int ay, int az, auto &reg) {
        constexpr int W = decltype(whd)::W, H = decltype(whd)::H, D = decltype(whd)::D;
        constexpr int NQ = (W * H * D + kMg3Block - 1) / kMg3Block;
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            const int q = tid + m * kMg3Block;
            const int zz = q / (W * H), rr = q - zz * W * H, yy = rr / W, xx = rr - yy * W;
            const int gx = ax + xx, gy = ay + yy, gz = az + zz;
            const bool ok = (q < W * H * D) & ((unsigned)gx < (unsigned)nx) & ((unsigned)gy < (unsigned)ny)
                            & ((unsigned)gz < (unsigned)nz);
            reg[m] = bufferLoad(b, ok ? 4 * (gx + nx * (gy + ny * gz)) : kNoOffset);
        }
    };
    auto store = [&](auto whd, const auto &reg, float *dst) {
        constexpr int W = decltype(whd)::W, H = decltype(whd)::H, D = decltype(whd)::D;
        constexpr int NQ = (W * H * D + kMg3Block - 1) / kMg3Block;
#pragma unroll
        for (int m = 0; m < NQ; ++m) {
            const int q = tid + m * kMg3Block;
            if (q < W * H * D) dst[q] = reg[m];
        }
    };
    using BX = Box<XW, XH, 4>;
    using BE = Box<FX + 1, FY, 2>;
    using BN = Box<FX, FY + 1, 2>;
    using BT = Box<FX, FY, 3>;
    using BS = Box<FX, FY, 2>;
    float rX[BX::NQ], rE[BE::NQ], rN[BN::NQ], rT[BT::NQ], rB[BS::NQ], rC[BS::NQ];
    load(BX{}, bX, fx0 - 1, fy0 - 1, fz0 - 1, rX);
    load(BE{}, bE, fx0 - 1, fy0, fz0, rE);
    load(BN{}, bN, fx0, fy0 - 1, fz0, rN);
    load(BT{}, bT, fx0, fy0, fz0 - 1, rT);
    load(BS{}, bB, fx0, fy0, fz0, rB);
    load(BS{}, bC, fx0, fy0, fz0, rC);
    store(BX{}, rX, &qx[0][0][0]);
    store(BE{}, rE, &qE[0][0][0]);
    store(BN{}, rN, &qN[0][0][0]);
    store(BT{}, rT, &qT[0][0][0]);
    store(BS{}, rB, &qB[0][0][0]);
    store(BS{}, rC, &qC[0][0][0]);
    __syncthreads();
    const int ti = tid % RX, tj = tid / RX;
    const int I = RX * bx + ti, J = RY * by + tj;
    if (I >= Cl.nx || J >= Cl.ny) return;
    double acc = 0.0;
    for (int dk = 0; dk < 2; ++dk) {
        const int k = fz0 + dk;
        if (k >= nz) break;
        for (int dj = 0; dj < 2; ++dj) {
            const int j = 2 * J + dj;
            if (j >= ny) break;
            for (int di = 0; di < 2; ++di) {
                const int i = 2 * I + di;
                if (i >= nx) break;
                const int li = 2 * ti + di, lj = 2 * tj + dj;          // in the block
                const float xc = qx[dk + 1][lj + 1][li + 1];
                double rr = (double)qB[dk][lj][li];
                rr -= (double)(qC[dk][lj][li] + mg3Outlet0(P, i, j, k, i + nx * (j + ny * k))) * (double)xc;
                rr -= (double)qE[dk][lj][li + 1] * (double)qx[dk + 1][lj + 1][li + 2];
                rr -= (double)qE[dk][lj][li] * (double)qx[dk + 1][lj + 1][li];
                rr -= (double)qN[dk][lj + 1][li] * (double)qx[dk + 1][lj + 2][li + 1];
                rr -= (double)qN[dk][lj][li] * (double)qx[dk + 1][lj][li + 1];
                rr -= (double)qT[dk + 1][lj][li] * (double)qx[dk + 2][lj + 1][li + 1];
                rr -= (double)qT[dk][lj][li] * (double)qx[dk][lj + 1][li + 1];
                acc += rr;
            }
        }
    }
    const int g = I + Cl.nx * (J + Cl.ny * K);
    Cl.b[g] = (float)acc;
    Cl.x[g] = 0.0f;
}

// ---------------- coarse correction (mgProlong) ----------------
template <bool L0>
__global__ __launch_bounds__(kMg3Block) void k_mg3_prolong(CaseParams P, MgLevel3 Fl, MgLevel3 Cl, float over) {
    if (caseDone(P)) return;
    const int g = blockIdx.x * kMg3Block + threadIdx.x;
    const Mg3View F = viewOf<L0>(P, Fl);
    if (g >= F.nx * F.ny * F.nz) return;
    const int i = g % F.nx, r = g / F.nx, j = r % F.ny, k = r / F.ny;
    const int p = i / 2 + Cl.nx * (j / 2 + Cl.ny * (k / 2));
    F.x[g] = F.x[g] + over * Cl.x[p];
}

// ---------------- one whole red-black sweep per launch, marching up z (large levels) ----------------
// A workgroup owns a MX x MY column over MZ planes and keeps a rolling window of four planes of x
// (2-cell halo in x and y) in LDS: at step t it relaxes the red cells of plane t+1 (tile plus a
// 1-cell ring), then the black cells of plane t (tile), then writes plane t.  red(t+1) reads black
// values of planes t..t+2, still old; black(t) reads red values of planes t-1..t+1, already new; the
// red planes below and above the chunk and the ring cells are recomputed from the same old values
// their owners use, so the launch equals the level-wide red pass followed by the black pass bit for
// bit.  It reads `in` and writes `out` (ping-pong).  The +x, +y and +z couplings come through LDS
// windows (a cell's -x, -y, -z couplings are its west, south and lower neighbours' +x, +y, +z ones);
// the source (and on coarse levels the stored centre) of each relaxed cell is read into registers
// one step ahead.  Level 0's centre is -(E + W + N + S) - (T + B) as k_pcorr_assemble forms it, plus
// the outlet term.  PROLONG: x is read as x + over * x_parent (mgProlong), the correction folded
// into the load of the first sweep after it.
constexpr int MX = 32, MY = 8, MZ = 32;

template <bool L0, bool PROLONG>
__global__ __launch_bounds__(kBlock) void k_mg3_zmarch(CaseParams P, MgLevel3 Fl, const float *in, float *out,
                                                       MgLevel3 Cl, float over) {
    constexpr int HX = MX + 4, HY = MY + 4, PL = HX * HY;
    constexpr int NPL = (PL + kBlock - 1) / kBlock;
    constexpr int RW = MX / 2 + 1, NR = (MY + 2) * RW;       // red cells of tile + ring per plane
    constexpr int BW = MX / 2, NB = MY * BW;                 // black cells of the tile per plane
    constexpr int NRT = (NR + kBlock - 1) / kBlock, NBT = (NB + kBlock - 1) / kBlock;
    __shared__ float win[4][HY][HX];                          // x, plane z in slot z & 3
    __shared__ float lT[4][HY][HX];                           // couplings towards +z, slot z & 3
    __shared__ float lE[3][HY][HX], lN[3][HY][HX];            // towards +x, +y, slot (z + 3) % 3
    if (caseDone(P)) return;
    const Mg3View F = viewOf<L0>(P, Fl);
    const int nx = F.nx, ny = F.ny, nz = F.nz, n = nx * ny * nz;
    const int gxc = (nx + MX - 1) / MX, gyc = (ny + MY - 1) / MY;
    const int w = xcdSwizzle((int)blockIdx.x, (int)gridDim.x);
    const int bx = w % gxc, byz = w / gxc;
    const int by = byz % gyc, bz = byz / gyc;
    const int x0 = bx * MX, y0 = by * MY, z0 = bz * MZ;
    const int z1 = min(z0 + MZ, nz);
    const int tid = threadIdx.x;
    const int nbytes = n * 4;
    const __amdgpu_buffer_rsrc_t bIn = bufferOf(in, nbytes), bSrc = bufferOf(F.b, nbytes);
    const __amdgpu_buffer_rsrc_t bE = bufferOf(F.E, nbytes), bN = bufferOf(F.N, nbytes), bT = bufferOf(F.T, nbytes);
    const __amdgpu_buffer_rsrc_t bC = bufferOf(F.C, nbytes);
    auto s3 = [](int z) { return (z + 3) % 3; };              // z >= -3

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
    // the old x of plane z, with the coarse correction added on the grid (PROLONG)
    auto loadX = [&](int z, float (&r)[NPL]) {
        loadPlane(bIn, z, r);
        if constexpr (PROLONG) {
#pragma unroll
            for (int m = 0; m < NPL; ++m) {
                const int q = tid + m * kBlock;
                const int ly = q / HX, lx = q - ly * HX;
                const int gx = x0 - 2 + lx, gy = y0 - 2 + ly;
                if ((q < PL) & ((unsigned)z < (unsigned)nz) & ((unsigned)gx < (unsigned)nx) & ((unsigned)gy < (unsigned)ny))
                    r[m] = r[m] + over * Cl.x[gx / 2 + Cl.nx * (gy / 2 + Cl.ny * (z / 2))];
            }
        }
    };
    auto storePlane = [&](float *wp, const float (&r)[NPL]) {
#pragma unroll
        for (int m = 0; m < NPL; ++m) {
            const int q = tid + m * kBlock;
            if (q < PL) wp[q] = r[m];
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
    // a relaxed cell's source and centre (level 0: the outlet term; coarse levels: the stored centre)
    struct Ops {
        float b, c;
    };
    auto loadOps = [&](bool ok, int gx, int gy, int z, Ops &o) {
        const int off = ok ? 4 * (gx + nx * (gy + ny * z)) : kNoOffset;
        o.b = bufferLoad(bSrc, off);
        if constexpr (L0) o.c = ok ? mg3Outlet0(P, gx, gy, z, gx + nx * (gy + ny * z)) : 0.0f;
        else o.c = bufferLoad(bC, off);
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
    auto writePlane = [&](int z) {
        for (int q = tid; q < MX * MY; q += kBlock) {
            const int r = q / MX, xx = q - r * MX;
            const int gx = x0 + xx, gy = y0 + r;
            if (gx < nx && gy < ny) out[gx + nx * (gy + ny * z)] = win[z & 3][r + 2][xx + 2];
        }
    };
    auto relax = [&](int gx, int gy, int z, const Ops &o) {
        const int lx = gx - x0 + 2, ly = gy - y0 + 2;
        const int c3 = s3(z);
        const float aE = lE[c3][ly][lx], aW = lE[c3][ly][lx - 1];
        const float aN = lN[c3][ly][lx], aS = lN[c3][ly - 1][lx];
        const float aT = lT[z & 3][ly][lx], aB = lT[(z - 1) & 3][ly][lx];
        float aC;
        if constexpr (L0) {
            aC = -(aE + aW + aN + aS);
            aC += -(aT + aB);
            aC = aC + o.c;
        } else {
            aC = o.c;
        }
        float(*wc)[HX] = win[z & 3];
        const float xE = wc[ly][lx + 1], xW = wc[ly][lx - 1];
        const float xN = wc[ly + 1][lx], xS = wc[ly - 1][lx];
        const float xT = (z != nz - 1) ? win[(z + 1) & 3][ly][lx] : 0.0f;
        const float xB = (z != 0) ? win[(z - 1) & 3][ly][lx] : 0.0f;
        float v = o.b - aE * xE - aW * xW - aN * xN - aS * xS;
        v -= aT * xT + aB * xB;
        wc[ly][lx] = v / aC;
    };
    auto redPlane = [&](int z, const Ops (&o)[NRT]) {
#pragma unroll
        for (int m = 0; m < NRT; ++m) {
            int gx, gy;
            if (redCell(z, m, gx, gy)) relax(gx, gy, z, o[m]);
        }
    };
    auto blackPlane = [&](int z, const Ops (&o)[NBT]) {
#pragma unroll
        for (int m = 0; m < NBT; ++m) {
            int gx, gy;
            if (blackCell(z, m, gx, gy)) relax(gx, gy, z, o[m]);
        }
    };

    float pend[NPL];
    Links pendL;
    Ops redB[NRT], blkB[NBT], redN[NRT], blkN[NBT];
    {
        float f[4][NPL];
        Links l[2];
        float tlow[NPL];
#pragma unroll
        for (int s = 0; s < 4; ++s) loadX(z0 - 2 + s, f[s]);
        loadPlane(bT, z0 - 2, tlow);
        loadLinks(z0 - 1, l[0]);
        loadLinks(z0, l[1]);
        loadX(z0 + 2, pend);
        loadLinks(z0 + 1, pendL);
        Ops r0[NRT];
        loadRed(z0 - 1, r0);
        loadRed(z0, redB);
        loadBlack(z0, blkB);
#pragma unroll
        for (int s = 0; s < 4; ++s) storePlane(&win[(z0 - 2 + s) & 3][0][0], f[s]);
        storePlane(&lT[(z0 - 2) & 3][0][0], tlow);
        storeLinks(z0 - 1, l[0]);
        storeLinks(z0, l[1]);
        __syncthreads();
        redPlane(z0 - 1, r0);
    }
    redPlane(z0, redB);
    loadRed(z0 + 1, redB);
    __syncthreads();
    for (int t = z0; t < z1; ++t) {
        storePlane(&win[(t + 2) & 3][0][0], pend);
        storeLinks(t + 1, pendL);
        loadX(t + 3, pend);
        loadLinks(t + 2, pendL);
        loadRed(t + 2, redN);
        loadBlack(t + 1, blkN);
        __syncthreads();
        redPlane(t + 1, redB);
        __syncthreads();
        blackPlane(t, blkB);
        __syncthreads();
        writePlane(t);
#pragma unroll
        for (int m = 0; m < NRT; ++m) redB[m] = redN[m];
#pragma unroll
        for (int m = 0; m < NBT; ++m) blkB[m] = blkN[m];
    }
}

// ---------------- the coarsest levels in one workgroup ----------------
// Levels lb .. nl-1 (at most kMg3BottomFloats of x, b and the matrix) of one V-cycle in LDS: down
// (NU sweeps, restriction), the coarsest level's sweeps, up (prolongation, NU sweeps); the same
// operations as the level-wide kernels, phase by phase between barriers.  Level lb's source comes
// restricted from level lb-1 and its correction goes back for the prolongation there.
constexpr int kMg3BottomThreads = 1024;

struct MgBottom3 {
    int lb, nl;
    MgLevel3 lv[kMgMaxLevels];
};

template <int NU>
__global__ __launch_bounds__(kMg3BottomThreads) void k_mg3_bottom(CaseParams P, MgBottom3 B, int ncoarse, float over) {
    __shared__ float lds[kMg3BottomFloats];
    if (caseDone(P)) return;
    const int tid = threadIdx.x;
    struct Lv {
        int nx, ny, nz, n;
        float *x, *b, *C, *E, *N, *T;
    } v[kMgMaxLevels];
    int off = 0;
    for (int m = B.lb; m < B.nl; ++m) {
        const MgLevel3 &F = B.lv[m];
        Lv &q = v[m];
        q.nx = F.nx;
        q.ny = F.ny;
        q.nz = F.nz;
        q.n = F.nx * F.ny * F.nz;
        float **pl[6] = {&q.x, &q.b, &q.C, &q.E, &q.N, &q.T};
        for (float **a : pl) {
            *a = lds + off;
            off += q.n;
        }
        for (int g = tid; g < q.n; g += kMg3BottomThreads) {
            q.x[g] = 0.0f;
            q.b[g] = m == B.lb ? F.b[g] : 0.0f;
            q.C[g] = F.C[g];
            q.E[g] = F.E[g];
            q.N[g] = F.N[g];
            q.T[g] = F.T[g];
        }
    }
    __syncthreads();
    auto sweeps = [&](const Lv &F, int n) {
        const int hx = (F.nx + 1) >> 1, pl = F.nx * F.ny;
        for (int s = 0; s < n; ++s)
            for (int col = 0; col < 2; ++col) {
                for (int g = tid; g < hx * F.ny * F.nz; g += kMg3BottomThreads) {
                    const int row = g / hx;
                    const int j = row % F.ny, k = row / F.ny;
                    const int i = 2 * (g - row * hx) + ((col + j + k) & 1);
                    if (i >= F.nx) continue;
                    const int c = i + F.nx * row;
                    const float *x = F.x;
                    const float xE = i < F.nx - 1 ? x[c + 1] : 0.0f, xW = i > 0 ? x[c - 1] : 0.0f;
                    const float xN = j < F.ny - 1 ? x[c + F.nx] : 0.0f, xS = j > 0 ? x[c - F.nx] : 0.0f;
                    const float aW = i > 0 ? F.E[c - 1] : 0.0f, aS = j > 0 ? F.N[c - F.nx] : 0.0f;
                    float val = F.b[c] - F.E[c] * xE - aW * xW - F.N[c] * xN - aS * xS;
                    const float xT = k < F.nz - 1 ? x[c + pl] : 0.0f, xB = k > 0 ? x[c - pl] : 0.0f;
                    const float aB = k > 0 ? F.T[c - pl] : 0.0f;
                    val -= F.T[c] * xT + aB * xB;
                    F.x[c] = val / F.C[c];
                }
                __syncthreads();
            }
    };
    auto res = [&](const Lv &F, int i, int j, int k) {
        const int pl = F.nx * F.ny, c = i + F.nx * (j + F.ny * k);
        const float *x = F.x;
        double r = (double)F.b[c];
        r -= (double)F.C[c] * (double)x[c];
        r -= (double)F.E[c] * (double)(i < F.nx - 1 ? x[c + 1] : 0.0f);
        r -= (double)(i > 0 ? F.E[c - 1] : 0.0f) * (double)(i > 0 ? x[c - 1] : 0.0f);
        r -= (double)F.N[c] * (double)(j < F.ny - 1 ? x[c + F.nx] : 0.0f);
        r -= (double)(j > 0 ? F.N[c - F.nx] : 0.0f) * (double)(j > 0 ? x[c - F.nx] : 0.0f);
        r -= (double)F.T[c] * (double)(k < F.nz - 1 ? x[c + pl] : 0.0f);
        r -= (double)(k > 0 ? F.T[c - pl] : 0.0f) * (double)(k > 0 ? x[c - pl] : 0.0f);
        return r;
    };
    for (int m = B.lb; m + 1 < B.nl; ++m) {
        const Lv &F = v[m], &Cl = v[m + 1];
        sweeps(F, NU);
        for (int g = tid; g < Cl.n; g += kMg3BottomThreads) {
            const int I = g % Cl.nx, R = g / Cl.nx, J = R % Cl.ny, K = R / Cl.ny;
            const int i1 = min(2 * I + 1, F.nx - 1), j1 = min(2 * J + 1, F.ny - 1), k1 = min(2 * K + 1, F.nz - 1);
            double acc = 0.0;
            for (int k = 2 * K; k <= k1; ++k)
                for (int j = 2 * J; j <= j1; ++j)
                    for (int i = 2 * I; i <= i1; ++i) acc += res(F, i, j, k);
            Cl.b[g] = (float)acc;
            Cl.x[g] = 0.0f;
        }
        __syncthreads();
    }
    sweeps(v[B.nl - 1], ncoarse);
    for (int m = B.nl - 2; m >= B.lb; --m) {
        const Lv &F = v[m], &Cl = v[m + 1];
        for (int g = tid; g < F.n; g += kMg3BottomThreads) {
            const int i = g % F.nx, r = g / F.nx, j = r % F.ny, k = r / F.ny;
            F.x[g] = F.x[g] + over * Cl.x[i / 2 + Cl.nx * (j / 2 + Cl.ny * (k / 2))];
        }
        __syncthreads();
        sweeps(F, NU);
    }
    const Lv &F = v[B.lb];
    for (int g = tid; g < F.n; g += kMg3BottomThreads) B.lv[B.lb].x[g] = F.x[g];
}

inline unsigned blocksFor(long long n) { return (unsigned)((n + kMg3Block - 1) / kMg3Block); }

// z-march columns of a level (the z-march pays once they fill the chip)
inline int zmarchColumns(const MgLevel3 &F) {
    return ((F.nx + MX - 1) / MX) * ((F.ny + MY - 1) / MY) * ((F.nz + MZ - 1) / MZ);
}
// (measured at 256^3: level 0 (2048 columns) 146 us per sweep against 2 x 99 us level-wide; level 1
// (256 columns) 36 us against 2 x 12.4 us)
inline bool useZmarch(const MgLevel3 &F) { return zmarchColumns(F) >= 512; }

// n sweeps of level l; prolong: add over * x_parent first (mgProlong).  Large levels sweep in
// z-march pairs (x -> partner -> x), the first folding in the correction; a single or odd last sweep
// runs level-wide in place.
void sweepLevel(const Lc &L, const MgLaunch3 &G, int l, int n, int &rec, bool prolong) {
    const MgLevel3 &F = G.lv[l];
    const MgLevel3 &Cl = G.lv[l + 1 < G.nl ? l + 1 : l];
    const unsigned nf = blocksFor((long long)F.nx * F.ny * F.nz);
    int s = 0;
    if (useZmarch(F) && n >= 2) {
        float *x = l == 0 ? L.P->pc : F.x, *xa = l == 0 ? L.P->alt[3] : F.xa;
        const unsigned nz = (unsigned)zmarchColumns(F);
        for (; s + 1 < n; s += 2) {
            if (l == 0) {
                if (prolong) k_mg3_zmarch<true, true><<<nz, kBlock, 0, L.s>>>(*L.P, F, x, xa, Cl, G.over);
                else k_mg3_zmarch<true, false><<<nz, kBlock, 0, L.s>>>(*L.P, F, x, xa, Cl, G.over);
                k_mg3_zmarch<true, false><<<nz, kBlock, 0, L.s>>>(*L.P, F, xa, x, Cl, G.over);
            } else {
                if (prolong) k_mg3_zmarch<false, true><<<nz, kBlock, 0, L.s>>>(*L.P, F, x, xa, Cl, G.over);
                else k_mg3_zmarch<false, false><<<nz, kBlock, 0, L.s>>>(*L.P, F, x, xa, Cl, G.over);
                k_mg3_zmarch<false, false><<<nz, kBlock, 0, L.s>>>(*L.P, F, xa, x, Cl, G.over);
            }
            prolong = false;
        }
    }
    if (prolong) {
        if (l == 0) k_mg3_prolong<true><<<nf, kMg3Block, 0, L.s>>>(*L.P, F, Cl, G.over);
        else k_mg3_prolong<false><<<nf, kMg3Block, 0, L.s>>>(*L.P, F, Cl, G.over);
    }
    const unsigned nb = blocksFor((long long)((F.nx + 1) >> 1) * F.ny * F.nz);
    for (; s < n; ++s)
        for (int col = 0; col < 2; ++col) {
            if (l == 0) k_mg3_sweep<true><<<nb, kMg3Block, 0, L.s>>>(*L.P, F, col, rec);
            else k_mg3_sweep<false><<<nb, kMg3Block, 0, L.s>>>(*L.P, F, col, rec);
            rec = 0;
        }
}

void launchBottom(const Lc &L, const MgLaunch3 &G) {
    MgBottom3 B{};
    B.lb = G.lb;
    B.nl = G.nl;
    for (int m = 0; m < G.nl; ++m) B.lv[m] = G.lv[m];
    if (G.nu == 1) k_mg3_bottom<1><<<1, kMg3BottomThreads, 0, L.s>>>(*L.P, B, G.ncoarse, G.over);
    else if (G.nu == 3) k_mg3_bottom<3><<<1, kMg3BottomThreads, 0, L.s>>>(*L.P, B, G.ncoarse, G.over);
    else k_mg3_bottom<2><<<1, kMg3BottomThreads, 0, L.s>>>(*L.P, B, G.ncoarse, G.over);
}

void cycleLevel(const Lc &L, const MgLaunch3 &G, int l, int &rec) {
    if (l == G.lb) {
        launchBottom(L, G);
        return;
    }
    if (l + 1 == G.nl) {
        sweepLevel(L, G, l, G.ncoarse, rec, false);
        return;
    }
    const MgLevel3 &F = G.lv[l], &Cl = G.lv[l + 1];
    sweepLevel(L, G, l, G.nu, rec, false);
    const unsigned nc = blocksFor((long long)Cl.nx * Cl.ny * Cl.nz);
    if (l == 0) {
        const unsigned nr = (unsigned)(((Cl.nx + RX - 1) / RX) * ((Cl.ny + RY - 1) / RY) * Cl.nz);
        k_mg3_restrict0<<<nr, kMg3Block, 0, L.s>>>(*L.P, F, Cl);
    } else {
        k_mg3_restrict<false><<<nc, kMg3Block, 0, L.s>>>(*L.P, F, Cl);
    }
    cycleLevel(L, G, l + 1, rec);
    sweepLevel(L, G, l, G.nu, rec, true);
}

}  // namespace

int mg3_levels(int nx, int ny, int nz, int (*dims)[3]) {
    int nl = 0;
    dims[0][0] = nx;
    dims[0][1] = ny;
    dims[0][2] = nz;
    ++nl;
    while (nl < kMgMaxLevels && dims[nl - 1][0] >= 4 && dims[nl - 1][1] >= 4 && dims[nl - 1][2] >= 4) {
        for (int a = 0; a < 3; ++a) dims[nl][a] = (dims[nl - 1][a] + 1) / 2;
        ++nl;
    }
    return nl;
}

int mg3_bottom_level(int nl, const int (*dims)[3]) {
    for (int l = 1; l < nl; ++l) {
        long long tot = 0;
        for (int m = l; m < nl; ++m) tot += 6LL * dims[m][0] * dims[m][1] * dims[m][2];
        if (tot <= kMg3BottomFloats) return l;
    }
    return nl;   // no level qualifies (level 0 alone, or a coarsest level too large): level-wide kernels
}

void launch_mg3_solve(const Lc &L, const MgLaunch3 &G) {
    int rec = G.cycles * 2 * G.nu;           // the sweep count for the record, written by the first launch
    for (int l = 0; l + 1 < G.nl; ++l) {
        const MgLevel3 &Cl = G.lv[l + 1];
        const unsigned nc = blocksFor((long long)Cl.nx * Cl.ny * Cl.nz);
        if (l == 0) k_mg3_coarsen<true><<<nc, kMg3Block, 0, L.s>>>(*L.P, G.lv[0], Cl, rec);
        else k_mg3_coarsen<false><<<nc, kMg3Block, 0, L.s>>>(*L.P, G.lv[l], Cl, 0);
        rec = 0;
    }
    for (int c = 0; c < G.cycles; ++c) cycleLevel(L, G, 0, rec);
}

}  // namespace ifl
