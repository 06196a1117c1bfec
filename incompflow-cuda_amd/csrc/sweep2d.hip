// sweep2d.hip - K (1, 2 or 4) full red-black Gauss-Seidel sweeps per launch for 2D grids.
//
// A workgroup owns a TX x TY tile.  It stages the old field on the tile plus a 2K-cell halo in
// LDS and runs the 2K half sweeps R1 B1 (R2 B2 ...) on shrinking regions: half sweep p updates
// the cells of its colour within 2K-1-p cells of the tile, which is exactly the set the later
// half sweeps read.  Cells outside the tile are recomputed from the same old values their owning
// tile uses, so they are bit-identical, and each launch equals the reference's red pass followed
// by its black pass, K times (GaussSeidelIterateKernel, kernels.cuh:926-1005, in red-black
// order).  Reads come from one buffer and writes go to the other (ping-pong), so no workgroup
// reads a value another workgroup has already updated.
//
//   * k_rb_tile2d: the momentum equations.  Coefficients are recomputed from the four face
//     velocities and the boundary rows (the functions of the assembly kernel, ifl_stencil.hpp),
//     one set for u and v, which are swept together in one launch since they share the matrix.
//     The temperature equation (COEF_TEMP) runs through the same kernel with its own diffusion
//     coefficient and boundary rows.
//   * k_rb_pc2d: the p' system, with the coefficients towards +x and +y of the stored p' matrix
//     (the coefficient towards -x of a cell is the +x one of its west neighbour, bit for bit).
// Every global load of the tile is issued first; the stop decision, which sums the previous
// launch's partials, is taken after the relax phases, with those loads in flight meanwhile.
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_stencil.hpp"

#include <utility>

namespace ifl {

// Colour-split LDS layout: cell (lx, ly) of a staged region has colour (lx + ly) & 1 (tile origins
// and halos are even, so this is the global red-black colour) and sits at [colour][ly][lx >> 1].
// The same-colour cells of a half sweep are then contiguous, and so are their neighbours.
template <int TX, int TY, int K, int NEQ, int KIND, bool BATCH, int NT>
__global__ __launch_bounds__(NT) void k_rb_tile2d(CaseParams P0, BatchMap M, SweepSel e0, SweepSel e1, int launch) {
    constexpr int H = 2 * K;                                   // field halo
    constexpr int R0 = 2 * K - 1;                              // ring of the first half sweep
    constexpr int HX = TX + 2 * H, HY = TY + 2 * H, HX2 = HX / 2;
    constexpr int SX = TX + 2 * R0, SY = TY + 2 * R0, SX2 = SX / 2;   // cells ever updated (sources)
    // the momentum and temperature matrices are rebuilt from the face velocities
    constexpr bool MOM = KIND == COEF_MOMENTUM || KIND == COEF_TEMP;
    static_assert(MOM, "the p' system has its own kernel (k_rb_pc2d)");
    constexpr int UX = SX + 1, UY = SY;                        // uf faces of the updated cells
    constexpr int VX = SX, VY = SY + 1;                        // vf faces of the updated cells
    constexpr int NST = (HX * HY + NT - 1) / NT;
    constexpr int NSR = (SX * SY + NT - 1) / NT;
    constexpr int NUT = MOM ? (UX * UY + NT - 1) / NT : 1;
    constexpr int NVT = MOM ? (VX * VY + NT - 1) / NT : 1;
    constexpr int NPF = 4;                                     // prefetched partials per thread and sweep
    static_assert(HX % 2 == 0 && SX % 2 == 0, "colour split needs even widths");
    __shared__ float sphi[NEQ][2][HY][HX2];
    __shared__ float ssrc[NEQ][2][SY][SX2];
    __shared__ float suf[MOM ? UY : 1][MOM ? UX : 1];
    __shared__ float svf[MOM ? VY : 1][MOM ? VX : 1];
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_TILE, blk);
    const int done = caseDone(P);          // checked after the loads are issued (they are in bounds)
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
        float *fld = slotField(P, eqs[q].slot), *alt = slotAlt(P, eqs[q].slot);
        const bool odd = ((launch + eqs[q].flip) & 1) != 0;
        in[q] = odd ? alt : fld;
        out[q] = odd ? fld : alt;
        src[q] = slotSource(P, eqs[q].slot);
        rlx[q] = P.relax[eqs[q].q];
    }
    const int tid = threadIdx.x;

    // ---- every global load of the tile first ----
    float rphi[NEQ][NST], rsrc[NEQ][NSR];
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int e = tid + m * NT;
        const int ly = e / HX, lx = e - ly * HX;
        const int gx = x0 - H + lx, gy = y0 - H + ly;
        const bool ok = e < HX * HY && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        const int g = gy * nx + gx;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) rphi[q][m] = ok ? in[q][g] : 0.0f;
    }
#pragma unroll
    for (int m = 0; m < NSR; ++m) {
        const int e = tid + m * NT;
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
            const int e = tid + m * NT;
            const int ly = e / UX, lx = e - ly * UX;
            const int fx = x0 - R0 + lx, gy = y0 - R0 + ly;
            const bool in = e < UX * UY && fx >= 0 && fx <= nx && gy >= 0 && gy < ny;
            ruf[m] = in ? P.uf[ufi(P, fx, gy, 0)] : 0.0f;
        }
#pragma unroll
        for (int m = 0; m < NVT; ++m) {
            const int e = tid + m * NT;
            const int ly = e / VX, lx = e - ly * VX;
            const int gx = x0 - R0 + lx, fy = y0 - R0 + ly;
            const bool in = e < VX * VY && gx >= 0 && gx < nx && fy >= 0 && fy <= ny;
            rvf[m] = in ? P.vf[vfi(P, gx, fy, 0)] : 0.0f;
        }
    }
    // ---- stage to LDS (colour split) ----
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int e = tid + m * NT;
        if (e < HX * HY) {
            const int ly = e / HX, lx = e - ly * HX;
            const int c = (lx + ly) & 1;
#pragma unroll
            for (int q = 0; q < NEQ; ++q) sphi[q][c][ly][lx >> 1] = rphi[q][m];
        }
    }
#pragma unroll
    for (int m = 0; m < NSR; ++m) {
        const int e = tid + m * NT;
        if (e < SX * SY) {
            const int ly = e / SX, lx = e - ly * SX;
            const int c = (lx + ly) & 1;
#pragma unroll
            for (int q = 0; q < NEQ; ++q) ssrc[q][c][ly][lx >> 1] = rsrc[q][m];
        }
    }
    if constexpr (MOM) {
#pragma unroll
        for (int m = 0; m < NUT; ++m) {
            const int e = tid + m * NT;
            if (e < UX * UY) suf[e / UX][e % UX] = ruf[m];
        }
#pragma unroll
        for (int m = 0; m < NVT; ++m) {
            const int e = tid + m * NT;
            if (e < VX * VY) svf[e / VX][e % VX] = rvf[m];
        }
    }
    // ---- partial sums of the previous launch's sweeps (the stop decision): the first NPF*256
    // records are loaded now (after the tile's registers are free) and added after the relax
    // phases, so their latency overlaps them ----
    const int nprev = blk.n;
    PrevState prev[NEQ];
#pragma unroll
    for (int q = 0; q < NEQ; ++q) prev[q] = loadPrev(P, eqs[q].slot, launch);
    double pf[NEQ][K][NPF];
#pragma unroll
    for (int q = 0; q < NEQ; ++q)
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const bool need = launch > 0 && kk < eqs[q].kPrev;
            const double *part = sweepPart(P, eqs[q].slot, launch - 1, kk);
#pragma unroll
            for (int f = 0; f < NPF; ++f) {
                const int bidx = tid + f * NT;
                pf[q][kk][f] = (need && bidx < nprev) ? part[bidx] : 0.0;
            }
        }
    if (done) return;                      // a finished case: nothing is written
    __syncthreads();

    double acc[K][2];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j][0] = acc[j][1] = 0.0;

    // relax the cell (gx, gy) of colour `col` in half sweep p (sweep p/2 of this launch)
    auto relax = [&](int gx, int gy, int col, int p) {
        const int lx = gx - x0 + H, ly = gy - y0 + H;     // field position
        const int ix = lx >> 1, ixE = (lx + 1) >> 1, ixW = (lx - 1) >> 1;
        const int oc = col ^ 1;
        float a[7];
        {
            const int sx = gx - x0 + R0, sy = gy - y0 + R0;
            const float uw = suf[sy][sx], ue = suf[sy][sx + 1];
            const float vs = svf[sy][sx], vn = svf[sy + 1][sx];
            if constexpr (KIND == COEF_TEMP) {
                linkCoef<2>(P, P.gam, ue, uw, vn, vs, 0.0f, 0.0f, a);
                tempBcRowsCoef<2>(P, gx, gy, 0, a);
            } else {
                momLinkCoef<2>(P, ue, uw, vn, vs, 0.0f, 0.0f, a);
                momBcRowsCoef<2>(P, gx, gy, 0, a);
            }
        }
        const bool owned = gx >= x0 && gx < x0 + TX && gy >= y0 && gy < y0 + TY;
        const int j = p >> 1;
        const int sxi = (lx - 1) >> 1, syi = ly - 1;       // source position (same colour)
#pragma unroll
        for (int q = 0; q < NEQ; ++q) {
            if (2 * eqs[q].k <= p) continue;
            const float phi_C = sphi[q][col][ly][ix];
            float newPhi = ssrc[q][col][syi][sxi] - a[sE] * sphi[q][oc][ly][ixE] - a[sW] * sphi[q][oc][ly][ixW]
                           - a[sN] * sphi[q][oc][ly + 1][ix] - a[sS] * sphi[q][oc][ly - 1][ix];
            newPhi /= a[sC];
            const float d = rlx[q] * (newPhi - phi_C);
            sphi[q][col][ly][ix] = phi_C + d;
            if (owned) {
#pragma unroll
                for (int jj = 0; jj < K; ++jj)
                    if (jj == j) acc[jj][q] = sqAdd(acc[jj][q], d);
            }
        }
    };

    const int kmax = eqs[0].k > eqs[NEQ - 1].k ? eqs[0].k : eqs[NEQ - 1].k;
    const int nph = 2 * kmax;
#pragma unroll
    for (int p = 0; p < 2 * K; ++p) {
        if (p >= nph) break;
        const int color = p & 1, ring = 2 * K - 1 - p;
        const int xs = x0 - ring, ys = y0 - ring;
        const int half = (TX + 2 * ring) / 2, rows = TY + 2 * ring;
        for (int e = tid; e < rows * half; e += NT) {
            const int r = e / half, kk = e - r * half;
            const int gy = ys + r;
            const int gx = xs + 2 * kk + ((xs + gy + color) & 1);
            if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) relax(gx, gy, color, p);
        }
        __syncthreads();
    }

    // ---- stop decision (launchRuns) from the prefetched and remaining partials ----
    bool act[2] = {true, true};
    {
        double t[NEQ * K];
#pragma unroll
        for (int q = 0; q < NEQ; ++q)
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                double s = 0.0;
                if (launch > 0 && kk < eqs[q].kPrev) {
                    const double *part = sweepPart(P, eqs[q].slot, launch - 1, kk);
#pragma unroll
                    for (int f = 0; f < NPF; ++f) s += pf[q][kk][f];
                    for (int bidx = tid + NPF * NT; bidx < nprev; bidx += NT) s += part[bidx];
                }
                t[q * K + kk] = s;
            }
        if (launch > 0) {
            double tot[NEQ * K];
            blockSum<NEQ * K, NT / 64>(t, tot);
#pragma unroll
            for (int q = 0; q < NEQ; ++q) {
                double tq[K];
#pragma unroll
                for (int kk = 0; kk < K; ++kk) tq[kk] = tot[q * K + kk];
                act[q] = eqs[q].k > 0 && launchDecide<K>(P, blk, eqs[q].slot, launch, eqs[q].kPrev, prev[q], tq, P.tol[eqs[q].q]);
            }
        } else {
#pragma unroll
            for (int q = 0; q < NEQ; ++q) act[q] = eqs[q].k > 0;
        }
    }

    // ---- write the tile: computed values, or the input unchanged for a stopped equation (keeps
    // the ping-pong parity) ----
    for (int e = tid; e < TX * TY; e += NT) {
        const int r = e / TX, xx = e - r * TX;
        const int gx = x0 + xx, gy = y0 + r;
        if (gx < nx && gy < ny) {
            const int g = gy * nx + gx;
            const int ly = r + H, lx = xx + H;
#pragma unroll
            for (int q = 0; q < NEQ; ++q) {
                const float val = act[q] ? sphi[q][(lx + ly) & 1][ly][lx >> 1] : in[q][g];
                out[q][g] = val;
                if (eqs[q].accum) P.pc[g] += val;       // the accumulation of k_pcorr_accumulate
            }
        }
    }
    // ---- this launch's sweep norms: one workgroup reduction for all equations and sweeps ----
    {
        double v[NEQ * K], tot[NEQ * K];
#pragma unroll
        for (int q = 0; q < NEQ; ++q)
#pragma unroll
            for (int j = 0; j < K; ++j) v[q * K + j] = acc[j][q];
        blockSum<NEQ * K, NT / 64>(v, tot);
#pragma unroll
        for (int q = 0; q < NEQ; ++q)
#pragma unroll
            for (int j = 0; j < K; ++j)     // constant indices: no private array in scratch
                if (tid == q * K + j && act[q] && j < eqs[q].k) sweepPart(P, eqs[q].slot, launch, j)[blk.flat] = tot[q * K + j];
    }
}

// ---------------- the p' system: a kernel of its own ----------------
// Same launch contract as k_rb_tile2d with one equation of kind COEF_PCORR, organised so that
// the relax loop is mostly arithmetic:
//   * one flat LDS block holds four planes with a common colour-split layout [colour][ly][ix]
//     over the field region: the field, the coefficients towards +x and +y, and the source
//     (rows 1..HY-2 only), so every operand of a cell is a compile-time offset from one of two
//     addresses (its own colour, the other colour);
//   * in half sweep p (compile-time) a thread owns one column ix of its colour and marches down
//     pairs of rows; the lx parity of a colour alternates row by row, so within a pair the E/W
//     offsets are constants;
//   * tiles whose ring lies inside the grid skip the per-cell grid test.
template <int TX, int TY, int K, bool BATCH, int NT>
__global__ __launch_bounds__(NT) void k_rb_pc2d(CaseParams P0, BatchMap M, SweepSel e0, int launch) {
    constexpr int H = 2 * K;
    constexpr int R0 = 2 * K - 1;
    constexpr int HX = TX + 2 * H, HY = TY + 2 * H, HX2 = HX / 2;
    constexpr int SX = TX + 2 * R0, SY = TY + 2 * R0;
    constexpr int NST = (HX * HY + NT - 1) / NT;
    constexpr int NSR = (SX * SY + NT - 1) / NT;
    constexpr int NPF = 4;
    constexpr int CO = HY * HX2;            // colour stride of the field and coefficient planes
    constexpr int PL = 2 * CO;              // their plane stride
    constexpr int CS = (HY - 2) * HX2;      // colour stride of the source plane (rows 1..HY-2)
    constexpr int oPHI = 0, oEX = PL, oEY = 2 * PL, oSRC = 3 * PL;
    static_assert(HX % 2 == 0, "colour split needs an even width");
    __shared__ float lds[3 * PL + 2 * CS];
    Blk blk;
    const CaseParams &P = caseOf<BATCH>(P0, M, MAP_TILE, blk);
    const int done = caseDone(P);
    const int nx = P.nx, ny = P.ny;
    const int tilesX = (nx + TX - 1) / TX;
    const int by = blk.flat / tilesX;
    const int bx = blk.flat - by * tilesX;
    const int x0 = bx * TX, y0 = by * TY;
    const SweepSel e = e0;
    float *fld = slotField(P, e.slot), *alt = slotAlt(P, e.slot);
    const bool oddL = ((launch + e.flip) & 1) != 0;
    const float *in = oddL ? alt : fld;
    float *out = oddL ? fld : alt;
    const float *src = slotSource(P, e.slot);
    const float rlx = P.relax[e.q];
    const int tid = threadIdx.x;

    // ---- global loads ----
    // the p' coefficients towards +x and +y are the stored planes of k_pcorr_assemble
    // (calcPresCorrLinkCoefKernel, kernels.cuh:517-529; 0 on the high boundary and off the grid)
    const float *cfE = P.pcoef + (size_t)P.N * sE, *cfN = P.pcoef + (size_t)P.N * sN;
    float rphi[NST], rex[NST], rey[NST], rsrc[NSR];
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int i = tid + m * NT;
        const int ly = i / HX, lx = i - ly * HX;
        const int gx = x0 - H + lx, gy = y0 - H + ly;
        const bool ok = i < HX * HY && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        const int g = gy * nx + gx;
        rphi[m] = ok ? in[g] : 0.0f;
        rex[m] = ok ? cfE[g] : 0.0f;
        rey[m] = ok ? cfN[g] : 0.0f;
    }
#pragma unroll
    for (int m = 0; m < NSR; ++m) {
        const int i = tid + m * NT;
        const int ly = i / SX, lx = i - ly * SX;
        const int gx = x0 - R0 + lx, gy = y0 - R0 + ly;
        const bool ok = i < SX * SY && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        rsrc[m] = ok ? src[gy * nx + gx] : 0.0f;
    }
    if (done) return;

    // ---- stage: field, coefficients, source at field coordinates ----
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int i = tid + m * NT;
        if (i < HX * HY) {
            const int ly = i / HX, lx = i - ly * HX;
            const int a = ((lx + ly) & 1) * CO + ly * HX2 + (lx >> 1);
            lds[oPHI + a] = rphi[m];
            lds[oEX + a] = rex[m];
            lds[oEY + a] = rey[m];
        }
    }
#pragma unroll
    for (int m = 0; m < NSR; ++m) {
        const int i = tid + m * NT;
        if (i < SX * SY) {
            const int sy = i / SX, sx = i - sy * SX;
            const int lx = sx + 1, ly = sy + 1;
            lds[oSRC + ((lx + ly) & 1) * CS + (ly - 1) * HX2 + (lx >> 1)] = rsrc[m];
        }
    }
    // partials of the previous launch (stop decision), in flight during the relax phases
    const int nprev = blk.n;
    const PrevState prev = loadPrev(P, e.slot, launch);
    double pf[K][NPF];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
        const bool need = launch > 0 && kk < e.kPrev;
        const double *part = sweepPart(P, e.slot, launch - 1, kk);
#pragma unroll
        for (int f = 0; f < NPF; ++f) {
            const int bidx = tid + f * NT;
            pf[kk][f] = (need && bidx < nprev) ? part[bidx] : 0.0;
        }
    }
    __syncthreads();
    // ---- the 2K half sweeps ----
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = 0.0;
    const int nph = 2 * e.k;
    auto phase = [&](auto pcst) {
        constexpr int p = decltype(pcst)::value;
        constexpr int col = p & 1, ring = 2 * K - 1 - p;
        constexpr int lyS = H - ring, rows = TY + 2 * ring;
        constexpr int lxS = H - ring, lxE = H + TX + ring;        // [lxS, lxE)
        constexpr int ixLo = lxS >> 1, ixHi = (lxE - 1) >> 1;
        constexpr int ncol = ixHi - ixLo + 1;
        constexpr int G = NT / ncol;                          // row-pair groups
        constexpr int npair = (rows + 1) / 2;
        constexpr int par0 = (lyS + col) & 1;                     // lx parity of the colour in row lyS
        static_assert(G >= 1, "tile too wide for one workgroup");
        // Lanes at the ends of a column range hold a cell one ring beyond this half sweep's region
        // and update it too: later half sweeps read only cells within their own ring plus one, and
        // it is not written out, so its value is never used.  Cells outside the grid are read as
        // neighbours (with coefficient 0) and must keep their staged 0: boundary tiles skip them.
        const bool interior = x0 - ring - 1 >= 0 && x0 + TX + ring + 1 <= nx && y0 - ring - 1 >= 0 && y0 + TY + ring + 1 <= ny;
        const int ci = tid % ncol, g = tid / ncol;
        if (g < G) {
            const int ix = ixLo + ci;
            for (int m = g; m < npair; m += G) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int ly = lyS + 2 * m + h;
                    const int par = h == 0 ? par0 : 1 - par0;
                    const int lx = 2 * ix + par;
                    const int gx = x0 - H + lx, gy = y0 - H + ly;
                    if (interior || (gx >= 0 && gx < nx && gy >= 0 && gy < ny)) {
                        const int own = col * CO + ly * HX2 + ix;
                        const int oth = (1 - col) * CO + ly * HX2 + ix;
                        const float aE = lds[oEX + own];
                        const float aW = lds[oEX + oth + par - 1];
                        const float aN = lds[oEY + own];
                        const float aS = lds[oEY + oth - HX2];
                        const float aC = -(aE + aW + aN + aS);
                        const float phi_C = lds[oPHI + own];
                        // source plane: own - col * (CO - CS) - HX2 (a constant offset)
                        float newPhi = lds[oSRC + own - col * (CO - CS) - HX2] - aE * lds[oPHI + oth + par]
                                       - aW * lds[oPHI + oth + par - 1]
                                       - aN * lds[oPHI + oth + HX2] - aS * lds[oPHI + oth - HX2];
                        newPhi /= aC;
                        const float d = rlx * (newPhi - phi_C);
                        lds[oPHI + own] = phi_C + d;
                        const bool owned = lx >= H && lx < H + TX && ly >= H && ly < H + TY;
                        acc[p >> 1] = sqAdd(acc[p >> 1], owned ? d : 0.0f);
                    }
                }
            }
        }
        __syncthreads();
    };
    [&]<int... Ps>(std::integer_sequence<int, Ps...>) {
        ((Ps < nph ? (phase(std::integral_constant<int, Ps>{}), 0) : 0), ...);
    }(std::make_integer_sequence<int, 2 * K>{});

    // ---- stop decision ----
    bool act = true;
    {
        double t[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            double sacc = 0.0;
            if (launch > 0 && kk < e.kPrev) {
                const double *part = sweepPart(P, e.slot, launch - 1, kk);
#pragma unroll
                for (int f = 0; f < NPF; ++f) sacc += pf[kk][f];
                for (int bidx = tid + NPF * NT; bidx < nprev; bidx += NT) sacc += part[bidx];
            }
            t[kk] = sacc;
        }
        if (launch > 0) {
            double tot[K];
            blockSum<K, NT / 64>(t, tot);
            act = e.k > 0 && launchDecide<K>(P, blk, e.slot, launch, e.kPrev, prev, tot, P.tol[e.q]);
        } else {
            act = e.k > 0;
        }
    }
    // ---- write the tile (or copy the input through for a stopped call) ----
    for (int i = tid; i < TX * TY; i += NT) {
        const int r = i / TX, xx = i - r * TX;
        const int gx = x0 + xx, gy = y0 + r;
        if (gx < nx && gy < ny) {
            const int g = gy * nx + gx;
            const int ly = r + H, lx = xx + H;
            const float val = act ? lds[oPHI + ((lx + ly) & 1) * CO + ly * HX2 + (lx >> 1)] : in[g];
            out[g] = val;
            if (e.accum) P.pc[g] += val;
        }
    }
    // ---- this launch's sweep norms ----
    {
        double tot[K];
        blockSum<K, NT / 64>(acc, tot);
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (tid == j && act && j < e.k) sweepPart(P, e.slot, launch, j)[blk.flat] = tot[j];
    }
}

template <int TX, int TY, int K, bool BATCH, int NT = kBlock>
static void launchTile(const Lc &L, int kind, const SweepSel &e0, const SweepSel &e1, int neq, int launch) {
    const CaseParams &P = *L.P;
    const dim3 grid = L.batch ? dim3((unsigned)L.total[MAP_TILE])
                              : dim3((unsigned)(((P.nx + TX - 1) / TX) * ((P.ny + TY - 1) / TY)));
    if (kind == COEF_PCORR) {
        k_rb_pc2d<TX, TY, K, BATCH, NT><<<grid, NT, 0, L.s>>>(P, L.M, e0, launch);
        return;
    }
    if (kind == COEF_TEMP) k_rb_tile2d<TX, TY, K, 1, COEF_TEMP, BATCH, NT><<<grid, NT, 0, L.s>>>(P, L.M, e0, e1, launch);
    else if (neq == 2) k_rb_tile2d<TX, TY, K, 2, COEF_MOMENTUM, BATCH, NT><<<grid, NT, 0, L.s>>>(P, L.M, e0, e1, launch);
    else k_rb_tile2d<TX, TY, K, 1, COEF_MOMENTUM, BATCH, NT><<<grid, NT, 0, L.s>>>(P, L.M, e0, e1, launch);
}

template <int K, bool BATCH>
static void launchK(const Lc &L, int kind, const SweepSel &e0, const SweepSel &e1, int neq, int launch) {
    switch (L.tile) {
        case TILE_128x32: launchTile<128, 32, K, BATCH, 1024>(L, kind, e0, e1, neq, launch); break;
        case TILE_64x16: launchTile<64, 16, K, BATCH>(L, kind, e0, e1, neq, launch); break;
        case TILE_64x8: launchTile<64, 8, K, BATCH>(L, kind, e0, e1, neq, launch); break;
        case TILE_32x8: launchTile<32, 8, K, BATCH>(L, kind, e0, e1, neq, launch); break;
        default: launchTile<32, 4, K, BATCH>(L, kind, e0, e1, neq, launch); break;
    }
}

static int tileX(int tile) { return tile == TILE_128x32 ? 128 : (tile == TILE_64x16 || tile == TILE_64x8) ? 64 : 32; }
static int tileY(int tile) { return tile == TILE_128x32 ? 32 : tile == TILE_64x16 ? 16 : tile == TILE_32x4 ? 4 : 8; }

int tile_dimx(int nx, int tile) { return (nx + tileX(tile) - 1) / tileX(tile); }
int tile_count(int nx, int ny, int tile) { return tile_dimx(nx, tile) * ((ny + tileY(tile) - 1) / tileY(tile)); }

// 1024^2 cavity, four p' sweeps per launch: 0.427 ms per outer iteration with 128x32 tiles,
// 0.486 ms with 64x16 (less halo recompute at the same occupancy: one 1024-thread workgroup or
// four 256-thread workgroups per CU)
int default_tile(int nx, int ny) {
    if (tile_count(nx, ny, TILE_128x32) >= 256) return TILE_128x32;
    if (tile_count(nx, ny, TILE_64x16) >= 768) return TILE_64x16;
    if (tile_count(nx, ny, TILE_64x8) >= 512) return TILE_64x8;
    if (tile_count(nx, ny, TILE_32x8) >= 256) return TILE_32x8;
    return TILE_32x4;
}

void launch_rb_sweep2d(const Lc &L, int kind, const SweepSel &e0, const SweepSel &e1, int neq, int launch, int block) {
    if (L.batch) {
        if (block == 4) launchK<4, true>(L, kind, e0, e1, neq, launch);
        else if (block >= 2) launchK<2, true>(L, kind, e0, e1, neq, launch);
        else launchK<1, true>(L, kind, e0, e1, neq, launch);
    } else {
        if (block == 4) launchK<4, false>(L, kind, e0, e1, neq, launch);
        else if (block >= 2) launchK<2, false>(L, kind, e0, e1, neq, launch);
        else launchK<1, false>(L, kind, e0, e1, neq, launch);
    }
}

}  // namespace ifl
