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
template <bool L0>
__global__ __launch_bounds__(kMg3Block) void k_mg3_coarsen(CaseParams P, MgLevel3 Fl, MgLevel3 Cl) {
    if (caseDone(P)) return;
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

inline unsigned blocksFor(long long n) { return (unsigned)((n + kMg3Block - 1) / kMg3Block); }

void sweepLevel(const Lc &L, const MgLaunch3 &G, int l, int n, int &rec) {
    const MgLevel3 &F = G.lv[l];
    const unsigned nb = blocksFor((long long)((F.nx + 1) >> 1) * F.ny * F.nz);
    for (int s = 0; s < n; ++s)
        for (int col = 0; col < 2; ++col) {
            if (l == 0) k_mg3_sweep<true><<<nb, kMg3Block, 0, L.s>>>(*L.P, F, col, rec);
            else k_mg3_sweep<false><<<nb, kMg3Block, 0, L.s>>>(*L.P, F, col, rec);
            rec = 0;
        }
}

void cycleLevel(const Lc &L, const MgLaunch3 &G, int l, int &rec) {
    if (l + 1 == G.nl) {
        sweepLevel(L, G, l, G.ncoarse, rec);
        return;
    }
    const MgLevel3 &F = G.lv[l], &Cl = G.lv[l + 1];
    sweepLevel(L, G, l, G.nu, rec);
    const unsigned nc = blocksFor((long long)Cl.nx * Cl.ny * Cl.nz);
    if (l == 0) k_mg3_restrict<true><<<nc, kMg3Block, 0, L.s>>>(*L.P, F, Cl);
    else k_mg3_restrict<false><<<nc, kMg3Block, 0, L.s>>>(*L.P, F, Cl);
    cycleLevel(L, G, l + 1, rec);
    const unsigned nf = blocksFor((long long)F.nx * F.ny * F.nz);
    if (l == 0) k_mg3_prolong<true><<<nf, kMg3Block, 0, L.s>>>(*L.P, F, Cl, G.over);
    else k_mg3_prolong<false><<<nf, kMg3Block, 0, L.s>>>(*L.P, F, Cl, G.over);
    sweepLevel(L, G, l, G.nu, rec);
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

void launch_mg3_solve(const Lc &L, const MgLaunch3 &G) {
    for (int l = 0; l + 1 < G.nl; ++l) {
        const MgLevel3 &Cl = G.lv[l + 1];
        const unsigned nc = blocksFor((long long)Cl.nx * Cl.ny * Cl.nz);
        if (l == 0) k_mg3_coarsen<true><<<nc, kMg3Block, 0, L.s>>>(*L.P, G.lv[0], Cl);
        else k_mg3_coarsen<false><<<nc, kMg3Block, 0, L.s>>>(*L.P, G.lv[l], Cl);
    }
    int rec = G.cycles * 2 * G.nu;
    for (int c = 0; c < G.cycles; ++c) cycleLevel(L, G, 0, rec);
}

}  // namespace ifl
