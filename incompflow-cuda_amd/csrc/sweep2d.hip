// sweep2d.hip - K (1, 2, 4 or 5) full red-black Gauss-Seidel sweeps per launch for 2D grids.
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
// One kernel, k_rb_sweep2d, serves the momentum equations (u and v together: they share the
// matrix), the temperature equation and the p' system.
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_ring.hpp"
#include "ifl_stencil.hpp"

#include <utility>

namespace ifl {

// K red-black sweeps of NEQ equations sharing one matrix (u and v; or T; or the p' system).
//   * The old fields are staged with a 2K halo in LDS, colour split: cell (lx, ly) has colour
//     (lx + ly) & 1 (tile origins and halos are even, so this is the global red-black colour) and
//     sits at [colour][ly][lx >> 1], so a cell's neighbours are the other colour's plane at
//     +-(row) and at two adjacent entries of its own row.
//   * A thread owns fixed cells of each colour (entries tid, tid + NT, ... of RingOrder's lists).
//     Their matrix rows (four links, the centre coefficient and the source), computed or read
//     once per launch, stay in registers through all 2K half sweeps; a half sweep touches LDS
//     only for the field.  Momentum and temperature rows come from the four face velocities and
//     the boundary rows (the functions of the assembly kernel, ifl_stencil.hpp); the p' rows are
//     the stored coefficients towards +x and +y (the coefficient towards -x of a cell is the +x
//     one of its west neighbour, bit for bit).
// Every global load of the tile is issued first; the stop decision, which sums the previous
// launch's partials, is taken after the relax phases, with those loads in flight meanwhile.
template <int TX, int TY, int K, int NEQ, int KIND, bool BATCH, int NT>
__global__ __launch_bounds__(NT) void k_rb_sweep2d(CaseParams P0, BatchMap M, SweepSel e0, SweepSel e1, int launch) {
    using RO = RingOrder<TX, TY, K>;
    constexpr bool PC = KIND == COEF_PCORR;
    static_assert(!PC || NEQ == 1, "the p' system is one equation");
    constexpr int H = 2 * K;                                   // field halo
    constexpr int R0 = 2 * K - 1;                              // ring of the first half sweep
    constexpr int HX = TX + 2 * H, HY = TY + 2 * H, HX2 = HX / 2;
    constexpr int SX = TX + 2 * R0, SY = TY + 2 * R0;          // cells ever relaxed
    constexpr int NST = (HX * HY + NT - 1) / NT;
    constexpr int NPF = NT >= 1024 ? 1 : 4;                    // prefetched partials per thread and sweep
    constexpr int NSL = (RO::NPC + NT - 1) / NT;               // cells per thread and colour
    constexpr int CO = HY * HX2;                               // colour stride of a field plane
    constexpr int PL = 2 * CO;                                 // field plane stride
    constexpr int UX = SX + 1, UY = SY, VX = SX, VY = SY + 1;  // face velocities of the relaxed cells
    constexpr int NUT = PC ? 1 : (UX * UY + NT - 1) / NT, NVT = PC ? 1 : (VX * VY + NT - 1) / NT;
    constexpr int CS = (HY - 2) * HX2;                         // colour stride of a source plane (rows 1..HY-2)
    constexpr int oSRC = NEQ * PL;                             // p' source plane
    constexpr int oEX = NEQ * PL + 2 * CS, oEY = oEX + PL;     // p' links towards +x, +y
    constexpr int NSR = PC ? (SX * SY + NT - 1) / NT : 1;      // staged p' sources per thread
    constexpr int oUF = NEQ * PL, oVF = oUF + UY * UX;         // momentum / temperature face planes
    constexpr int NLDS = PC ? oEX + 2 * PL : oVF + VY * VX;
    // byte offsets, from a cell's base address, of its other-colour north/south neighbours and
    // of the cell itself (red cells: the base is the cell; black cells: its south neighbour)
    constexpr int bN0 = 4 * (CO + HX2), bS0 = 4 * (CO - HX2), bN1 = 4 * 2 * HX2, bO1 = 4 * (CO + HX2);
    static_assert(4 * (NEQ - 1) * PL + bN0 < 65536, "LDS immediate offsets are 16 bits");
    static_assert(HX % 2 == 0, "colour split needs an even width");
    __shared__ float lds[NLDS];
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

    // ---- this thread's cells (entries tid, tid + NT, ... of each colour's list; 0xffff: none),
    // looked up first: the face velocity loads and the row setup wait for them ----
    int cell[2][NSL];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int sl = 0; sl < NSL; ++sl) {
            const int i = sl * NT + tid;
            cell[c][sl] = ((sl + 1) * NT <= RO::NPC || i < RO::NPC) ? RO::table.cell[c][i] : 0xffff;
        }

    // ---- every global load of the tile first: the matrix data, then the field ----
    // Raw buffer loads: one descriptor per array (N floats) and a 32-bit byte offset per cell (a
    // case holds at most 2^28 cells, so 4 N < 2^31); a cell outside the grid gets an offset past
    // the array, which the hardware reads as 0, so no branch surrounds a load.
    const int nbytes = P.N * 4;
    __amdgpu_buffer_rsrc_t bIn[NEQ], bSrc[NEQ];
#pragma unroll
    for (int q = 0; q < NEQ; ++q) {
        bIn[q] = bufferOf(in[q], nbytes);
        bSrc[q] = bufferOf(src[q], nbytes);
    }
    // matrix rows in registers: the four links, the centre coefficient and the source of each cell
    float cE[2][NSL], cW[2][NSL], cN[2][NSL], cS[2][NSL], cC[2][NSL], cB[NEQ][2][NSL];
    float rex[PC ? NST : 1], rey[PC ? NST : 1], rsrc[NSR], ruf[NUT], rvf[NVT];
    if constexpr (PC) {
        // the stored links towards +x and +y of k_pcorr_assemble (calcPresCorrLinkCoefKernel,
        // kernels.cuh:517-529; 0 on the high boundary and off the grid) over the staged region, the
        // source over the relaxed cells
        const float *cfE = P.pcoef + (size_t)P.N * sE, *cfN = P.pcoef + (size_t)P.N * sN;
        const __amdgpu_buffer_rsrc_t bE = bufferOf(cfE, nbytes), bN = bufferOf(cfN, nbytes);
#pragma unroll
        for (int m = 0; m < NST; ++m) {
            const int i = tid + m * NT;
            const int ly = i / HX, lx = i - ly * HX;
            const int gx = x0 - H + lx, gy = y0 - H + ly;
            const bool ok = (i < HX * HY) & ((unsigned)gx < (unsigned)nx) & ((unsigned)gy < (unsigned)ny);
            const int off = ok ? 4 * (gy * nx + gx) : kNoOffset;
            rex[m] = bufferLoad(bE, off);
            rey[m] = bufferLoad(bN, off);
        }
#pragma unroll
        for (int m = 0; m < NSR; ++m) {
            const int i = tid + m * NT;
            const int ly = i / SX, lx = i - ly * SX;
            const int gx = x0 - R0 + lx, gy = y0 - R0 + ly;
            const bool ok = (i < SX * SY) & ((unsigned)gx < (unsigned)nx) & ((unsigned)gy < (unsigned)ny);
            rsrc[m] = bufferLoad(bSrc[0], ok ? 4 * (gy * nx + gx) : kNoOffset);
        }
    } else {
        // face velocities over the relaxed cells, coalesced, for LDS; each cell's sources straight
        // into registers (off-grid faces and cells read as 0)
        const __amdgpu_buffer_rsrc_t bUf = bufferOf(P.uf, 4 * (nx + 1) * ny), bVf = bufferOf(P.vf, 4 * nx * (ny + 1));
#pragma unroll
        for (int m = 0; m < NUT; ++m) {
            const int i = tid + m * NT;
            const int ly = i / UX, lx = i - ly * UX;
            const int fx = x0 - R0 + lx, gy = y0 - R0 + ly;
            const bool ok = (i < UX * UY) & ((unsigned)fx <= (unsigned)nx) & ((unsigned)gy < (unsigned)ny);
            ruf[m] = bufferLoad(bUf, ok ? 4 * ufi(P, fx, gy, 0) : kNoOffset);
        }
#pragma unroll
        for (int m = 0; m < NVT; ++m) {
            const int i = tid + m * NT;
            const int ly = i / VX, lx = i - ly * VX;
            const int gx = x0 - R0 + lx, fy = y0 - R0 + ly;
            const bool ok = (i < VX * VY) & ((unsigned)gx < (unsigned)nx) & ((unsigned)fy <= (unsigned)ny);
            rvf[m] = bufferLoad(bVf, ok ? 4 * vfi(P, gx, fy, 0) : kNoOffset);
        }
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int sl = 0; sl < NSL; ++sl) {
                const int e = cell[c][sl];
                const int gx = x0 - H + (e & 255), gy = y0 - H + (e >> 8);
                const bool ok = (e != 0xffff) & ((unsigned)gx < (unsigned)nx) & ((unsigned)gy < (unsigned)ny);
                const int off = ok ? 4 * (gy * nx + gx) : kNoOffset;
#pragma unroll
                for (int q = 0; q < NEQ; ++q) cB[q][c][sl] = bufferLoad(bSrc[q], off);
            }
    }
    // the field last
    float rphi[NEQ][NST];
#pragma unroll
    for (int m = 0; m < NST; ++m) {
        const int i = tid + m * NT;
        const int ly = i / HX, lx = i - ly * HX;
        const int gx = x0 - H + lx, gy = y0 - H + ly;
        // one unsigned compare per axis, no short circuit: no branch around the offset either
        const bool ok = (i < HX * HY) & ((unsigned)gx < (unsigned)nx) & ((unsigned)gy < (unsigned)ny);
        const int off = ok ? 4 * (gy * nx + gx) : kNoOffset;
#pragma unroll
        for (int q = 0; q < NEQ; ++q) rphi[q][m] = bufferLoad(bIn[q], off);
    }
    // ---- stage the matrix data to LDS ----
    if constexpr (PC) {
#pragma unroll
        for (int m = 0; m < NST; ++m) {
            const int i = tid + m * NT;
            if (i < HX * HY) {
                const int ly = i / HX, lx = i - ly * HX;
                const int a = ((lx + ly) & 1) * CO + ly * HX2 + (lx >> 1);
                lds[oEX + a] = rex[m];
                lds[oEY + a] = rey[m];
            }
        }
#pragma unroll
        for (int m = 0; m < NSR; ++m) {
            const int i = tid + m * NT;
            if (i < SX * SY) {
                const int sy = i / SX, sx = i - sy * SX;
                const int lx = sx + 1, ly = sy + 1;            // field coordinates (R0 = H - 1)
                const int a = ((lx + ly) & 1) * CS + (ly - 1) * HX2 + (lx >> 1);
                lds[oSRC + a] = rsrc[m];
            }
        }
    } else {
        // the face velocities of the momentum and temperature rows
#pragma unroll
        for (int m = 0; m < NUT; ++m) {
            const int i = tid + m * NT;
            if (i < UX * UY) lds[oUF + i] = ruf[m];
        }
#pragma unroll
        for (int m = 0; m < NVT; ++m) {
            const int i = tid + m * NT;
            if (i < VX * VY) lds[oVF + i] = rvf[m];
        }
    }
    // one equation: the field is staged after the row setup, its loads overlapping that work; two
    // (u and v) stage it now, which keeps their registers for the rows
    constexpr bool kLateField = NEQ == 1;
    auto stageField = [&]() {
#pragma unroll
        for (int m = 0; m < NST; ++m) {
            const int i = tid + m * NT;
            if (i < HX * HY) {
                const int ly = i / HX, lx = i - ly * HX;
                const int a = ((lx + ly) & 1) * CO + ly * HX2 + (lx >> 1);
#pragma unroll
                for (int q = 0; q < NEQ; ++q) lds[q * PL + a] = rphi[q][m];
            }
        }
    };
    if constexpr (!kLateField) stageField();
    // ---- partial sums of the previous launch's sweeps (the stop decision): the first NPF*NT
    // records are loaded now and added after the relax phases, so their latency overlaps them ----
    const int nprev = blk.n;
    PrevState prev[NEQ];
#pragma unroll
    for (int q = 0; q < NEQ; ++q) prev[q] = loadPrev(P, eqs[q].slot, launch);
    // (read unconditionally, at record 0 where there is none: a conditional load would be sunk to
    // its use after the relax phases; launch 0 reads the other parity's records, also in bounds)
    double pf[NEQ][K][NPF];
#pragma unroll
    for (int q = 0; q < NEQ; ++q)
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const double *part = sweepPart(P, eqs[q].slot, launch - 1, kk);
#pragma unroll
            for (int f = 0; f < NPF; ++f) {
                const int bidx = tid + f * NT;
                pf[q][kk][f] = part[bidx < nprev ? bidx : 0];
            }
        }
    if (done) return;                      // a finished case: nothing is written
    __syncthreads();

    // ---- this thread's cells: LDS addresses and matrix rows ----
    // ab: the byte address of the cell's base in its equation's field plane (red cells: the cell;
    // black cells: its south neighbour in the red plane), -1 for none or outside the grid; wb: the
    // byte address of its west neighbour in the other colour's plane (the east neighbour follows).
    int ab[2][NSL], wb[2][NSL];
    const LinkConsts lk = linkConsts<2>(P, KIND == COEF_TEMP ? P.gam : P.mu);
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int s = 0; s < NSL; ++s) {
            const int i = s * NT + tid;
            ab[c][s] = -1;
            wb[c][s] = 0;
            cE[c][s] = cW[c][s] = cN[c][s] = cS[c][s] = 0.0f;
            cC[c][s] = 1.0f;
            if constexpr (PC) cB[0][c][s] = 0.0f;
            if ((s + 1) * NT <= RO::NPC || i < RO::NPC) {
                const int e = cell[c][s];
                const int lx = e & 255, ly = e >> 8;
                const int gx = x0 - H + lx, gy = y0 - H + ly;
                if (((unsigned)gx < (unsigned)nx) & ((unsigned)gy < (unsigned)ny)) {
                    const int o = c * CO + ly * HX2 + (lx >> 1);
                    const int ot = (1 - c) * CO + ly * HX2 + (lx >> 1);
                    const int w = ot + (lx & 1) - 1;
                    ab[c][s] = 4 * (c == 0 ? o : o - CO - HX2);
                    wb[c][s] = 4 * w;
                    if constexpr (PC) {
                        cB[0][c][s] = lds[oSRC + o - c * (CO - CS) - HX2];
                        const float aE = lds[oEX + o], aW = lds[oEX + w];
                        const float aN = lds[oEY + o], aS = lds[oEY + ot - HX2];
                        cE[c][s] = aE;
                        cW[c][s] = aW;
                        cN[c][s] = aN;
                        cS[c][s] = aS;
                        cC[c][s] = -(aE + aW + aN + aS);
                    } else {
                        const int sx = lx - 1, sy = ly - 1;    // face coordinates (R0 = H - 1)
                        const float uw = lds[oUF + sy * UX + sx], ue = lds[oUF + sy * UX + sx + 1];
                        const float vs = lds[oVF + sy * VX + sx], vn = lds[oVF + (sy + 1) * VX + sx];
                        float a[7];
                        if constexpr (KIND == COEF_TEMP) {
                            linkCoef<2>(lk, ue, uw, vn, vs, 0.0f, 0.0f, a);
                            tempBcRowsCoef<2>(P, gx, gy, 0, a);
                        } else {
                            linkCoef<2>(lk, ue, uw, vn, vs, 0.0f, 0.0f, a);
                            momBcRowsCoef<2>(P, gx, gy, 0, a);
                        }
                        cE[c][s] = a[sE];
                        cW[c][s] = a[sW];
                        cN[c][s] = a[sN];
                        cS[c][s] = a[sS];
                        cC[c][s] = a[sC];
                    }
                }
            }
        }
    // ---- stage the field (its own LDS planes: the rows above read only the matrix planes) ----
    if constexpr (kLateField) stageField();
    __syncthreads();

    // ---- the 2K half sweeps ----
    double acc[K][NEQ];
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
        for (int q = 0; q < NEQ; ++q) acc[j][q] = 0.0;
    const int kmax = eqs[0].k > eqs[NEQ - 1].k ? eqs[0].k : eqs[NEQ - 1].k;
    const int nph = 2 * kmax;
    auto phase = [&](auto pcst) {
        constexpr int p = decltype(pcst)::value;
        constexpr int col = p & 1, ring = 2 * K - 1 - p;
        constexpr int cnt = RO::upTo(ring), core = RO::upTo(0);
        // byte offsets from the base: the cell itself, its north and south neighbours
        constexpr int bO = col == 0 ? 0 : bO1, bN = col == 0 ? bN0 : bN1, bS = col == 0 ? bS0 : 0;
#pragma unroll
        for (int s = 0; s < NSL; ++s) {
            if (s * NT < cnt) {          // constant after unrolling
                const int i = s * NT + tid;
                const int a = ab[col][s];
                if (a >= 0 && ((s + 1) * NT <= cnt || i < cnt)) {
                    const int w = wb[col][s];
                    const bool isCore = (s + 1) * NT <= core || i < core;
#pragma unroll
                    for (int q = 0; q < NEQ; ++q) {
                        if (2 * eqs[q].k <= p) continue;
                        char *f = (char *)(lds + q * PL);
                        const float phi_C = *(const float *)(f + a + bO);
                        const float fE = *(const float *)(f + w + 4), fW = *(const float *)(f + w);
                        const float fN = *(const float *)(f + a + bN), fS = *(const float *)(f + a + bS);
                        float newPhi = cB[q][col][s] - cE[col][s] * fE - cW[col][s] * fW - cN[col][s] * fN
                                       - cS[col][s] * fS;
                        newPhi /= cC[col][s];
                        const float d = rlx[q] * (newPhi - phi_C);
                        *(float *)(f + a + bO) = phi_C + d;
                        if (isCore) acc[p >> 1][q] = sqAdd(acc[p >> 1][q], d);
                    }
                }
            }
        }
        __syncthreads();
    };
    [&]<int... Ps>(std::integer_sequence<int, Ps...>) {
        ((Ps < nph ? (phase(std::integral_constant<int, Ps>{}), 0) : 0), ...);
    }(std::make_integer_sequence<int, 2 * K>{});

    // ---- the workgroup sums: the stop decision's (launchRuns: the previous launch's partials,
    // prefetched and remaining) and this launch's sweep norms, each total finished in a wave of its
    // own.  Every wave adds: the previous launch's records past nprev are +0, which leaves the sums
    // as they are. ----
    constexpr int NW = NT / 64, ND = NEQ * K, NR = (ND + NW - 1) / NW;
    const int wave = tid >> 6, lane = tid & 63;
    bool act[2] = {true, true};
    double totD[NR], totN[NR];      // tot[r]: the total of value wave + r NW
    __shared__ double sTot[ND];
    __shared__ int sAct[2];
    if (launch > 0) {
        double t[ND];
#pragma unroll
        for (int q = 0; q < NEQ; ++q)
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                double sacc = 0.0;
                if (kk < eqs[q].kPrev) {
                    const double *part = sweepPart(P, eqs[q].slot, launch - 1, kk);
#pragma unroll
                    for (int f = 0; f < NPF; ++f) sacc += (tid + f * NT < nprev) ? pf[q][kk][f] : 0.0;
                    for (int bidx = tid + NPF * NT; bidx < nprev; bidx += NT) sacc += part[bidx];
                }
                t[q * K + kk] = sacc;
            }
        blockSumSpread<ND, NW, 0>(t, totD);
        // the decision's totals go to wave 0 (read after the next reduction's barrier)
#pragma unroll
        for (int s = 0; s < ND; ++s)
            if (s % NW == wave && wave != 0 && lane == 0) sTot[s] = totD[s / NW];
    }
    {
        double v[ND];
#pragma unroll
        for (int q = 0; q < NEQ; ++q)
#pragma unroll
            for (int j = 0; j < K; ++j) v[q * K + j] = acc[j][q];
        blockSumSpread<ND, NW, 1>(v, totN);
    }
    if (launch > 0) {
        // wave 0's lanes take the decision's FP64 square roots and divisions one sweep per lane
        // (the leader, thread 0, publishes the state); the decision reaches the other waves through LDS
        if (wave == 0) {
#pragma unroll
            for (int q = 0; q < NEQ; ++q) {
                double tq[K];
#pragma unroll
                for (int kk = 0; kk < K; ++kk) tq[kk] = (q * K + kk) % NW == 0 ? totD[(q * K + kk) / NW] : sTot[q * K + kk];
                const bool a = eqs[q].k > 0 && launchDecideWave<K>(P, blk, eqs[q].slot, launch, eqs[q].kPrev, prev[q], tq, P.tol[eqs[q].q]);
                if (tid == 0) sAct[q] = a ? 1 : 0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NEQ; ++q) act[q] = sAct[q] != 0;
    } else {
#pragma unroll
        for (int q = 0; q < NEQ; ++q) act[q] = eqs[q].k > 0;
    }
    // this launch's partials, each stored by the wave that finished it
#pragma unroll
    for (int q = 0; q < NEQ; ++q)
#pragma unroll
        for (int j = 0; j < K; ++j) {   // constant indices: no private array in scratch
            const int sIdx = q * K + j;
            if (sIdx % NW == wave && lane == 0 && act[q] && j < eqs[q].k)
                sweepPart(P, eqs[q].slot, launch, j)[blk.flat] = totN[sIdx / NW];
        }

    // ---- write the tile: computed values, or the input unchanged for a stopped equation (keeps
    // the ping-pong parity) ----
    for (int i = tid; i < TX * TY; i += NT) {
        const int r = i / TX, xx = i - r * TX;
        const int gx = x0 + xx, gy = y0 + r;
        if (gx < nx && gy < ny) {
            const int g = gy * nx + gx;
            const int ly = r + H, lx = xx + H;
#pragma unroll
            for (int q = 0; q < NEQ; ++q) {
                const float val = act[q] ? lds[q * PL + ((lx + ly) & 1) * CO + ly * HX2 + (lx >> 1)] : in[q][g];
                out[q][g] = val;
                if (eqs[q].accum) P.pc[g] += val;       // the accumulation of k_pcorr_accumulate
            }
        }
    }
}

template <int TX, int TY, int K, bool BATCH, int NT = kBlock>
static void launchTile(const Lc &L, int kind, const SweepSel &e0, const SweepSel &e1, int neq, int launch) {
    const CaseParams &P = *L.P;
    const dim3 grid = L.batch ? dim3((unsigned)L.total[MAP_TILE])
                              : dim3((unsigned)(((P.nx + TX - 1) / TX) * ((P.ny + TY - 1) / TY)));
    if (kind == COEF_PCORR) k_rb_sweep2d<TX, TY, K, 1, COEF_PCORR, BATCH, NT><<<grid, NT, 0, L.s>>>(P, L.M, e0, e1, launch);
    else if (kind == COEF_TEMP) k_rb_sweep2d<TX, TY, K, 1, COEF_TEMP, BATCH, NT><<<grid, NT, 0, L.s>>>(P, L.M, e0, e1, launch);
    else if (neq == 2) k_rb_sweep2d<TX, TY, K, 2, COEF_MOMENTUM, BATCH, NT><<<grid, NT, 0, L.s>>>(P, L.M, e0, e1, launch);
    else k_rb_sweep2d<TX, TY, K, 1, COEF_MOMENTUM, BATCH, NT><<<grid, NT, 0, L.s>>>(P, L.M, e0, e1, launch);
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
        if (block == 5) launchK<5, true>(L, kind, e0, e1, neq, launch);
        else if (block == 4) launchK<4, true>(L, kind, e0, e1, neq, launch);
        else if (block >= 2) launchK<2, true>(L, kind, e0, e1, neq, launch);
        else launchK<1, true>(L, kind, e0, e1, neq, launch);
    } else {
        if (block == 5) launchK<5, false>(L, kind, e0, e1, neq, launch);
        else if (block == 4) launchK<4, false>(L, kind, e0, e1, neq, launch);
        else if (block >= 2) launchK<2, false>(L, kind, e0, e1, neq, launch);
        else launchK<1, false>(L, kind, e0, e1, neq, launch);
    }
}

}  // namespace ifl
