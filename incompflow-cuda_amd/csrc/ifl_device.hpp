// ifl_device.hpp - device-side data structures and helpers of the gfx950 SIMPLE solver.
//
// One case (grid) is described by CaseParams (passed by value to every kernel) and owns one
// CaseState in device memory that carries the FP64 reduction results, the smoother stop state
// and the outer-iteration monitor.
//
// Reductions never synchronise workgroups inside a kernel (an agent-scope fence per workgroup
// writes back the XCD's L2 on gfx950).  Each workgroup stores its FP64 partial sums with plain
// stores; a later kernel (after the kernel boundary has made them visible) adds them up in a fixed
// order, so every total is deterministic:
//   * smoother sweeps: every workgroup of sweep s re-adds the partials of sweep s-1 to decide,
//     identically in all workgroups, whether sweep s still runs (Solver.cu:955-962);
//   * assembly/update sums: one single-workgroup kernel per outer iteration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/incompflow.h"

namespace ifl {

constexpr int kBlock = 256;            // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kSlots = 6;              // smoother call slots: u, v, w, p', p' refinement, temperature
constexpr int kSlotT = 5;              // the temperature slot (its equation index is 4)
constexpr int kRows = 4;               // most rows per thread in the cell kernels (64 x 16 cells per workgroup)

enum { E = 0, W = 1, N = 2, S = 3, T = 4, B = 5 };                // constants.h:7-12
enum { sC = 0, sE = 1, sW = 2, sN = 3, sS = 4, sT = 5, sB = 6 };  // constants.h:14-20

// State of one smoother call after launch L (a launch runs 1 or 2 sweeps), double-buffered by
// the parity of L so that the workgroups of launch L+1 can read the state after launch L-1 while
// workgroup 0 publishes the state after L.
struct SmoothState {
    double maxNorm[2];  // running max of the sweep RMS increment (Solver.cu:941, 957)
    int stopped[2];     // relNorm < tol seen (the remaining sweeps of the call are skipped)
    int swAt[2];        // sweeps that ran up to and including that launch
    double lastSum;     // sum of squared increments of the last sweep that ran
    int sweeps;         // sweeps that ran in the current call
    int pad;
};

struct CaseState {
    int done;               // 0 running, 1 converged, 2 diverged, 3 iteration cap: every kernel returns at once
    int pad0;
    long long iter;         // outer iterations executed
    SmoothState sm[kSlots];
    int sweepsIter[5];      // sweeps executed in the current outer iteration (u, v, w, p', T)
    double mass;            // sum b'^2 (calcPresCorrSrcTerm)
    double resSq[4];        // sum (b - A phi)^2 of u, v, w and T before the sweeps
    double upd[7];          // sums of squared increments p,u,v,w,uf,vf,wf (updateField)
    double pNormMax, uNormMax, vNormMax, wNormMax, ufNormMax, vfNormMax, wfNormMax;
    double resMax[5];       // scaled-residual normalisers: continuity, u, v, w, T
    double lastScaled;
    int lastConverged;
    int pad;
    long long iterCap;      // > 0: stop (done 3) once iter reaches it (batched solves)
};

struct CaseParams {
    int dim, nx, ny, nz;
    int N;                   // cells (< 2^28, checked on the host)
    int cellRows;            // rows per thread of the cell kernels (1 or kRows): 64 x 4 cellRows cells per workgroup
    int zcM, zcP;            // z planes per column chunk of the 3D z-march sweeps (momentum/temperature, p')
    float dx, dy, dz, Ax, Ay, Az, rho, mu;
    int bct[6];
    float bcv[6][3];
    const float *prof[6];
    float relax[5];          // u, v, w, p', T
    double tol[5];
    double tolMass, tolRes;
    int stopRule;
    int histCap;
    int energy;              // the temperature equation is solved (its buffers below exist)
    float gam;               // k / c_p: the diffusion coefficient of the temperature equation
    int tbct[6];             // temperature BC per face: IFL_TEMP_DIRICHLET / IFL_TEMP_NEUMANN
    float tbcv[6];           // Dirichlet values
    float *tcoef, *bt;       // temperature matrix ((1+2dim) planes) and source
    float *u, *v, *w, *p, *pc, *uf, *vf, *wf, *temp;
    float *mcoef, *pcoef, *bu, *bv, *bw, *bp;
    float *tmp, *tmp2;       // p' defect and refinement correction
    float *alt[6];           // ping-pong partners of u, v, w, p', refinement correction, T
    float *rinv;             // 1/aC of the momentum matrix (p' coefficients are rebuilt from it)
    double *spart;           // sweep partials: [slot][launch parity][sweep in launch][partStride]
    double *partMom;         // 3 per cell workgroup: momentum residual sums
    double *partMass;        // 1 per cell workgroup: b'^2
    double *partUpd;         // 7 per cell workgroup: update increments
    int partStride;          // doubles per (slot, parity) sweep-partial array
    CaseState *st;
    ifl_iter_record *hist;   // ring of histCap records indexed by iter % histCap
};

// ---------------- indexing (kernels.cuh:30-38, 105-114) ----------------
__device__ __forceinline__ int cid(const CaseParams &P, int i, int j, int k) { return i + P.nx * (j + P.ny * k); }
__device__ __forceinline__ int ufi(const CaseParams &P, int i, int j, int k) { return i + (P.nx + 1) * (j + P.ny * k); }
__device__ __forceinline__ int vfi(const CaseParams &P, int i, int j, int k) { return i + P.nx * (j + (P.ny + 1) * k); }
__device__ __forceinline__ int wfi(const CaseParams &P, int i, int j, int k) { return i + P.nx * (j + P.ny * k); }

// velBCs::val[f][comp] (config.h:35-42) or the per-face profile of the normal component.
__device__ __forceinline__ float bcval(const CaseParams &P, int f, int comp, int q) {
    if (comp == (f >> 1) && P.prof[f] != nullptr) return P.prof[f][q];
    return P.bcv[f][comp];
}
// Boundary face velocity, kernels.cuh:87-93: wall 0, inlet value, outlet adjacent cell.
__device__ __forceinline__ float faceBC(const CaseParams &P, int f, int q, float adjacent) {
    int t = P.bct[f];
    if (t == IFL_WALL) return 0.0f;
    if (t == IFL_INLET) return bcval(P, f, f >> 1, q);
    return adjacent;
}
__device__ __forceinline__ bool ghostCopy(const CaseParams &P, int f) {
    return P.bct[f] == IFL_WALL || P.bct[f] == IFL_INLET;
}
// std::min(x, 0.0f) with its exact selection rule ((0 < x) ? 0 : x).
__device__ __forceinline__ float min0(float x) { return (0.0f < x) ? 0.0f : x; }

// Raw buffer access (gfx9 descriptor: stride 0, dword 3 = 0x00020000): loads at a byte offset at
// or past num_records return 0.
constexpr int kNoOffset = (int)0x80000000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bufferOf(const float *p, int nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, nbytes, 0x00020000);
}
__device__ __forceinline__ float bufferLoad(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
// (A case holds at most 2^28 cells, so the byte offsets of its fields fit in 31 bits.)

// ---------------- FP64 reductions (fixed order) ----------------
// acc + d * d in FP64 as one fused operation: the square of a float is exact in FP64, so this
// equals the separate multiply and add bit for bit.
__device__ __forceinline__ double sqAdd(double acc, float d) {
    const double x = (double)d;
    return __builtin_fma(x, x, acc);
}
// A double moved across lanes by a DPP pattern (both 32-bit halves); every pattern used here is a
// permutation inside each row of 16 lanes, so every lane reads an active lane of its own row.
template <int CTRL>
__device__ __forceinline__ double dppMove(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double laneValue(double x, int lane) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// A double from the previous row's lane 15 (CTRL row_bcast:15) or from lane 31 (row_bcast:31) in
// the rows of ROWMASK, 0 in the others.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dppBcast(double x) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWMASK, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
// K sums over each row of 16 lanes at once, all lanes active: pairs (quad_perm 1,0,3,2), quads
// (quad_perm 2,3,0,1), half rows (row_half_mirror) and rows (row_mirror) are added, each step
// commutative, so every lane of a row ends with the same row total.  The K chains are
// interleaved step by step, so their latencies overlap.
template <int K>
__device__ __forceinline__ void rowSums(double (&x)[K]) {
#pragma unroll
    for (int s = 0; s < K; ++s) x[s] += dppMove<0xB1>(x[s]);
#pragma unroll
    for (int s = 0; s < K; ++s) x[s] += dppMove<0x4E>(x[s]);
#pragma unroll
    for (int s = 0; s < K; ++s) x[s] += dppMove<0x141>(x[s]);
#pragma unroll
    for (int s = 0; s < K; ++s) x[s] += dppMove<0x140>(x[s]);
}
// K wave64 sums, the totals landing in lane 63: row sums, then row_bcast:15 adds R0 to row 1 and
// R2 to row 3, and row_bcast:31 adds row 1's R1 + R0 to row 3: lane 63 holds (R3 + R2) + (R1 + R0).
template <int K>
__device__ __forceinline__ void waveSums(double (&x)[K]) {
    rowSums<K>(x);
#pragma unroll
    for (int s = 0; s < K; ++s) x[s] += dppBcast<0x142, 0xA>(x[s]);
#pragma unroll
    for (int s = 0; s < K; ++s) x[s] += dppBcast<0x143, 0xC>(x[s]);
}

// Reduce K (<= 64) per-thread values over the workgroup of NW (<= 16) waves (all threads must
// call); the lanes of wave 0 return the K totals (the other waves' out is unspecified).  Each
// wave adds its rows of 16 lanes and stores the four row totals; wave 0 then loads row total l
// into lane l and adds its 64 lanes.  Waves from `active` on hold only zeros and skip their part
// (a sum of non-negative terms does not change when zeros are left out).  A second call of the
// same instantiation may follow only after a barrier that wave 0 reaches after reading the first.
template <int K, int NW = kWaves>
__device__ __forceinline__ void blockSumWave0(const double (&v)[K], double (&out)[K], int active = NW) {
    static_assert(K <= 64 && NW <= 16, "one value per lane, 64 row totals");
    __shared__ double lds[4 * NW][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave < active) {
        double w[K];
#pragma unroll
        for (int s = 0; s < K; ++s) w[s] = v[s];
        rowSums<K>(w);
        if ((lane & 15) == 15)
#pragma unroll
            for (int s = 0; s < K; ++s) lds[4 * wave + (lane >> 4)][s] = w[s];
    }
    __syncthreads();
    if (wave == 0) {
        double r[K];
        const bool has = lane < 4 * active;
#pragma unroll
        for (int s = 0; s < K; ++s) r[s] = has ? lds[lane][s] : 0.0;
        waveSums<K>(r);
#pragma unroll
        for (int s = 0; s < K; ++s) out[s] = laneValue(r[s], 63);
    }
}

// The sums of blockSumWave0 (every wave adding: the four row totals of each wave, then a wave64
// sum of the 4 NW row totals in lane order), with value s finished by wave s % NW, so the final
// wave sums of different values run in different waves.  Every lane of wave w receives in out[r]
// the total of value w + r NW (where that is < K).  v is used as scratch.  All threads call.
// Instantiations with different TAGs use LDS arrays of their own, so they may follow each other
// without a barrier between them.
template <int K, int NW, int TAG = 0>
__device__ __forceinline__ void blockSumSpread(double (&v)[K], double (&out)[(K + NW - 1) / NW]) {
    static_assert(NW <= 16, "64 row totals");
    __shared__ double lds[4 * NW][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    rowSums<K>(v);
    if ((lane & 15) == 15)
#pragma unroll
        for (int s = 0; s < K; ++s) lds[4 * wave + (lane >> 4)][s] = v[s];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < (K + NW - 1) / NW; ++r) {
        const int s = wave + r * NW;
        double x[1] = {(s < K && lane < 4 * NW) ? lds[lane][s < K ? s : 0] : 0.0};
        waveSums<1>(x);
        out[r] = laneValue(x[0], 63);
    }
}

// The same with every thread returning the K totals.
template <int K, int NW = kWaves>
__device__ __forceinline__ void blockSum(const double (&v)[K], double (&out)[K]) {
    static_assert(K <= 64, "one value per lane");
    __shared__ double bc[K];
    double t[K];
    blockSumWave0<K, NW>(v, t);
    if (threadIdx.x < K) {
        double mine = t[0];
#pragma unroll
        for (int s = 1; s < K; ++s) mine = (int)threadIdx.x == s ? t[s] : mine;
        bc[threadIdx.x] = mine;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < K; ++s) out[s] = bc[s];
    __syncthreads();
}

// Store this workgroup's K totals at out[0..K).
template <int K>
__device__ __forceinline__ void storeBlockSums(const double (&v)[K], double *out) {
    double tot[K];
    blockSumWave0<K>(v, tot);
#pragma unroll
    for (int s = 0; s < K; ++s)
        if ((int)threadIdx.x == s) out[s] = tot[s];
}

// Totals of n partial records of K doubles (record b at part[b*K]); thread t adds records
// t, t+256, ... in order, then the fixed workgroup tree: the same result in every workgroup.
template <int K>
__device__ __forceinline__ void sumPartials(const double *part, int n, double (&out)[K]) {
    double acc[K];
#pragma unroll
    for (int s = 0; s < K; ++s) acc[s] = 0.0;
    for (int b = threadIdx.x; b < n; b += kBlock)
#pragma unroll
        for (int s = 0; s < K; ++s) acc[s] += part[b * K + s];
    blockSum<K>(acc, out);
}

__device__ __forceinline__ int caseDone(const CaseParams &P) { return P.st->done; }

// ---------------- one case or a batch of cases per launch ----------------
// A kernel runs either on one case (its CaseParams is a kernel argument and the grid is the
// case's own) or on a batch: the grid is the concatenation of every case's workgroups, and a
// workgroup finds its case in a per-workgroup table.  Either way the kernel body sees the case's
// parameters and the workgroup's position inside that case's own grid.
// MAP_ZMARCH: 3D momentum/temperature z-march columns; MAP_ZMARCH_P: 3D p' columns (own shape)
enum { MAP_CELL = 0, MAP_TILE = 1, MAP_ONE = 2, MAP_COPY = 3, MAP_ZMARCH = 4, MAP_ZMARCH_P = 5, MAP_KINDS = 6 };

struct BatchMap {
    int ncase;
    const CaseParams *P;            // device array, one per case
    const int *start[MAP_KINDS];    // prefix sums of the workgroups per case (ncase + 1)
    const int *dimx[MAP_KINDS];     // x extent of each case's grid
    const int4 *blk[MAP_KINDS];     // per workgroup: {case, index in the case's grid, its size, x extent}
};

struct Blk {
    int x, y, z;      // workgroup position in its case's grid
    int gx, gy, gz;   // that grid's extents
    int flat;         // x + gx (y + gy z)
    int n;            // workgroups in the case's grid
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup w to XCD w % 8), and each XCD
// has its own L2.  The work item of workgroup w is chosen so that each XCD receives one contiguous
// range of work items (tiles, columns, rows): neighbouring tiles then share their halos and
// z-neighbour planes in one L2 instead of fetching them twice from the Infinity Fabric.  A
// bijection of [0, n): XCD x gets the n/8 (+1 for the first n%8 XCDs) items from x*(n/8)+min(x,n%8).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcdSwizzle(int w, int n) {
    const int x = w % kXcds, i = w / kXcds;
    const int q = n / kXcds, r = n % kXcds;
    return x * q + min(x, r) + i;
}

template <bool BATCH>
__device__ __forceinline__ const CaseParams &caseOf(const CaseParams &P0, const BatchMap &M, int kind, Blk &b) {
    if constexpr (BATCH) {
        // one (scalar) load: the workgroup's entry of the batch's block table
        const int4 e = M.blk[kind][xcdSwizzle((int)blockIdx.x, (int)gridDim.x)];
        const int c = __builtin_amdgcn_readfirstlane(e.x);
        const int loc = __builtin_amdgcn_readfirstlane(e.y);
        const int n = __builtin_amdgcn_readfirstlane(e.z);
        const int gx = __builtin_amdgcn_readfirstlane(e.w);
        b.gx = gx;
        b.gy = n / gx;
        b.gz = 1;
        b.x = loc % gx;
        b.y = loc / gx;
        b.z = 0;
        b.flat = loc;
        b.n = n;
        return M.P[c];
    } else {
        b.gx = gridDim.x; b.gy = gridDim.y; b.gz = gridDim.z;
        b.n = b.gx * b.gy * b.gz;
        const int w = (int)(blockIdx.x + b.gx * (blockIdx.y + b.gy * blockIdx.z));
        // the 3D cell kernels (64x16-cell rows, no z reuse inside an XCD's L2 at this size) measured
        // faster in dispatch order: 256^3 assembly 651 us against 682 us swizzled
        const int t = (kind == MAP_CELL && P0.dim == 3) ? w : xcdSwizzle(w, b.n);
        b.flat = t;
        b.x = t % b.gx;
        b.y = (t / b.gx) % b.gy;
        b.z = t / (b.gx * b.gy);
        return P0;
    }
}

// Workgroups of `kind` of the case (batch) or nSingle (one case).
template <bool BATCH>
__device__ __forceinline__ int kindCount(const BatchMap &M, int kind, const CaseParams &P, int nSingle) {
    if constexpr (BATCH) {
        const int c = (int)(&P - M.P);
        return M.start[kind][c + 1] - M.start[kind][c];
    } else {
        return nSingle;
    }
}

__device__ __forceinline__ bool leader(const Blk &b) { return b.flat == 0 && threadIdx.x == 0; }

// The field a smoother slot works on (u, v, w, p', refinement correction, T) and its source.
__device__ __forceinline__ float *slotField(const CaseParams &P, int slot) {
    return slot == 0 ? P.u : slot == 1 ? P.v : slot == 2 ? P.w : slot == 3 ? P.pc : slot == 4 ? P.tmp2 : P.temp;
}
// The slot's ping-pong partner (a select chain: indexing the kernel-argument array with a
// runtime value would copy the whole argument block to scratch memory).
__device__ __forceinline__ float *slotAlt(const CaseParams &P, int slot) {
    return slot == 0 ? P.alt[0] : slot == 1 ? P.alt[1] : slot == 2 ? P.alt[2] : slot == 3 ? P.alt[3]
         : slot == 4 ? P.alt[4] : P.alt[5];
}
__device__ __forceinline__ const float *slotSource(const CaseParams &P, int slot) {
    return slot == 0 ? P.bu : slot == 1 ? P.bv : slot == 2 ? P.bw : slot == 3 ? P.bp : slot == 4 ? P.tmp : P.bt;
}
__device__ __forceinline__ const float *slotCoef(const CaseParams &P, int slot) {
    return slot < 3 ? P.mcoef : slot == kSlotT ? P.tcoef : P.pcoef;
}

// Sweep partials: [slot][launch parity][sweep in the launch, < kMaxBlock][partStride].
constexpr int kMaxBlock = 5;    // most red-black sweeps per launch
__device__ __forceinline__ double *sweepPart(const CaseParams &P, int slot, int launch, int kk) {
    return P.spart + (size_t)((slot * 2 + (launch & 1)) * kMaxBlock + kk) * P.partStride;
}

// Whether launch L of the smoother call in `slot` runs.  Every workgroup evaluates, from their
// partials (nprev records each), the kPrev sweeps of launch L-1 in order exactly as the host loop
// of Solver.cu:955-962 does after each sweep (in FP64): norm, running max, relative norm against
// tol.  A met test stops the call after the launch that contains the sweep (with one sweep per
// launch: right after that sweep, as in the reference).  Workgroup 0 publishes the state after
// launch L-1.  With L == number of launches this closes the call (sweeps count, last sum).
// The stop state after launch L-2 of a smoother call (what launchRuns reads), loaded early by
// kernels that overlap these loads with their own work.
struct PrevState {
    bool stoppedBefore;
    double mx;
    int before;
};
__device__ __forceinline__ PrevState loadPrev(const CaseParams &P, int slot, int L) {
    PrevState v{false, -1e20, 0};
    if (L >= 2) {
        const SmoothState &c = P.st->sm[slot];
        const int pp = L & 1;
        v.stoppedBefore = c.stopped[pp] != 0;
        v.mx = c.maxNorm[pp];
        v.before = c.swAt[pp];
    }
    return v;
}

// launchRuns with the totals T[kk] of the previous launch's sweeps already reduced by the caller
// and the state already loaded.
template <int K>
__device__ __forceinline__ bool launchDecide(const CaseParams &P, const Blk &b, int slot, int L, int kPrev,
                                             const PrevState &pv, const double (&T)[K], double tol) {
    if (L == 0) return true;
    double mx = pv.mx;
    bool stopped = pv.stoppedBefore;
    double last = 0.0;
    int ran = pv.before;
    if (!pv.stoppedBefore) {
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            if (kk < kPrev) {
                last = T[kk];
                const double norm = sqrt(last / (double)P.N);
                mx = fmax(norm, mx);
                if (norm / (mx + 1e-20) < tol) stopped = true;
            }
        }
        ran = pv.before + kPrev;
    }
    if (leader(b)) {
        SmoothState &c = P.st->sm[slot];
        const int q = (L - 1) & 1;
        c.stopped[q] = stopped ? 1 : 0;
        c.maxNorm[q] = mx;
        c.swAt[q] = ran;
        if (!pv.stoppedBefore) c.lastSum = last;
        c.sweeps = ran;
    }
    return !stopped;
}

// launchDecide for wave 0 of a workgroup (all its lanes): lane kk computes sweep kk's norm, so the
// K square roots and divisions run side by side.  The values and the decision are launchDecide's
// bit for bit: each lane takes the running maximum over the sweeps up to its own in the loop's
// order (fmax is exact), and the call stops if any sweep's test is met.
template <int K>
__device__ __forceinline__ bool launchDecideWave(const CaseParams &P, const Blk &b, int slot, int L, int kPrev,
                                                 const PrevState &pv, const double (&T)[K], double tol) {
    if (L == 0) return true;
    static_assert(K <= 64, "one sweep per lane");
    const int lane = threadIdx.x & 63;
    double mx = pv.mx;
    bool stopped = pv.stoppedBefore;
    double last = 0.0;
    int ran = pv.before;
    if (!pv.stoppedBefore) {
        double tk = T[0];
#pragma unroll
        for (int kk = 1; kk < K; ++kk) tk = lane == kk ? T[kk] : tk;
        const double norm = sqrt(tk / (double)P.N);
        double m = pv.mx;
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const double nk = laneValue(norm, kk);
            if (kk < kPrev && kk <= lane) m = fmax(nk, m);
        }
        const bool hit = lane < kPrev && norm / (m + 1e-20) < tol;
        stopped = __builtin_amdgcn_ballot_w64(hit) != 0;
        if (kPrev > 0) mx = laneValue(m, kPrev - 1);
#pragma unroll
        for (int kk = 0; kk < K; ++kk) last = kk == kPrev - 1 ? T[kk] : last;
        ran = pv.before + kPrev;
    }
    if (leader(b)) {
        SmoothState &c = P.st->sm[slot];
        const int q = (L - 1) & 1;
        c.stopped[q] = stopped ? 1 : 0;
        c.maxNorm[q] = mx;
        c.swAt[q] = ran;
        if (!pv.stoppedBefore) c.lastSum = last;
        c.sweeps = ran;
    }
    return !stopped;
}

// Smoother calls whose closing step (k_smooth_finish) workgroup 0 of another kernel performs.
struct Finish {
    int n = 0;
    int kind = 0;      // workgroup kind of the calls' launches (their partial counts)
    int nprev = 0;     // their workgroup count (one case)
    int slot[4] = {0, 0, 0, 0};
    int q[4] = {0, 0, 0, 0};
    int nL[4] = {0, 0, 0, 0};
    int kLast[4] = {0, 0, 0, 0};
};

__device__ __forceinline__ bool launchRuns(const CaseParams &P, const Blk &b, int slot, int L, int kPrev, int nprev,
                                           double tol) {
    if (L == 0) return true;
    SmoothState &c = P.st->sm[slot];
    const int pp = L & 1;                          // parity of the state after launch L-2
    const bool stoppedBefore = (L >= 2) && c.stopped[pp];
    double mx = (L >= 2) ? c.maxNorm[pp] : -1e20;
    const int before = (L >= 2) ? c.swAt[pp] : 0;
    bool stopped = stoppedBefore;
    double T = 0.0;
    int ran = before;
    if (!stoppedBefore) {
        for (int kk = 0; kk < kPrev; ++kk) {
            double t[1];
            sumPartials<1>(sweepPart(P, slot, L - 1, kk), nprev, t);
            T = t[0];
            const double norm = sqrt(T / (double)P.N);
            mx = fmax(norm, mx);
            if (norm / (mx + 1e-20) < tol) stopped = true;
        }
        ran = before + kPrev;
    }
    if (leader(b)) {
        const int q = (L - 1) & 1;
        c.stopped[q] = stopped ? 1 : 0;
        c.maxNorm[q] = mx;
        c.swAt[q] = ran;
        if (!stoppedBefore) c.lastSum = T;
        c.sweeps = ran;
    }
    return !stopped;
}

// Close the smoother calls listed in fin: publish the state after each call's last launch (the
// decision of launchRuns for a launch past the last) and add its sweeps to the outer iteration's
// count.  Each call is closed by one workgroup of the case, which sums the
// partials of all the launch's sweeps in one workgroup reduction (the same per-sweep totals as
// one reduction per sweep), so the closing adds one short reduction to one workgroup's path.  The
// workgroup is chosen by the call's equation (q modulo the workgroup count), so calls that add to
// the same sweep count (p' and its refinement correction) are closed by one workgroup, in order.
template <bool BATCH>
__device__ __forceinline__ void runFinish(const CaseParams &P, const BatchMap &M, const Blk &b, const Finish &fin) {
    if (fin.n == 0 || b.flat >= 5) return;
    int np = fin.nprev;
    if constexpr (BATCH) np = M.start[fin.kind][(int)(&P - M.P) + 1] - M.start[fin.kind][(int)(&P - M.P)];
    for (int i = 0; i < fin.n; ++i) {
        if (fin.q[i] % b.n != b.flat) continue;
        const int slot = fin.slot[i], L = fin.nL[i], kPrev = fin.kLast[i];
        const double tol = P.tol[fin.q[i]];
        SmoothState &c = P.st->sm[slot];
        const int pp = L & 1;
        const bool stoppedBefore = (L >= 2) && c.stopped[pp];
        double mx = (L >= 2) ? c.maxNorm[pp] : -1e20;
        const int before = (L >= 2) ? c.swAt[pp] : 0;
        bool stopped = stoppedBefore;
        double T = 0.0;
        int ran = before;
        if (L >= 1 && !stoppedBefore) {
            double acc[kMaxBlock], tot[kMaxBlock];
#pragma unroll
            for (int kk = 0; kk < kMaxBlock; ++kk) acc[kk] = 0.0;
            for (int r = threadIdx.x; r < np; r += kBlock)
#pragma unroll
                for (int kk = 0; kk < kMaxBlock; ++kk)
                    if (kk < kPrev) acc[kk] += sweepPart(P, slot, L - 1, kk)[r];
            blockSum<kMaxBlock>(acc, tot);
#pragma unroll
            for (int kk = 0; kk < kMaxBlock; ++kk)
                if (kk < kPrev) {
                    T = tot[kk];
                    const double norm = sqrt(T / (double)P.N);
                    mx = fmax(norm, mx);
                    if (norm / (mx + 1e-20) < tol) stopped = true;
                }
            ran = before + kPrev;
        }
        if (threadIdx.x == 0 && L >= 1) {
            const int q = (L - 1) & 1;
            c.stopped[q] = stopped ? 1 : 0;
            c.maxNorm[q] = mx;
            c.swAt[q] = ran;
            if (!stoppedBefore) c.lastSum = T;
            c.sweeps = ran;
        }
        if (threadIdx.x == 0) P.st->sweepsIter[fin.q[i]] += L >= 1 ? ran : c.sweeps;
    }
}

}  // namespace ifl
