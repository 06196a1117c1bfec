// ifl_device.hpp - device-side data structures and helpers of the gfx950 SIMPLE solver.
//
// One case (grid) is described by CaseParams (passed by value to every kernel, or read from a
// device table in the batched path) and owns one CaseState in device memory that carries the
// FP64 reduction results, the smoother stop flags and the outer-iteration monitor.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/incompflow.h"

namespace ifl {

constexpr int kBlock = 256;            // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kPart = 8;               // FP64 partial sums per workgroup
constexpr int kSlots = 5;              // smoother call slots: u, v, w, p', p' refinement

enum { E = 0, W = 1, N = 2, S = 3, T = 4, B = 5 };                // constants.h:7-12
enum { sC = 0, sE = 1, sW = 2, sN = 3, sS = 4, sT = 5, sB = 6 };  // constants.h:14-20

struct SmoothState {
    double maxNorm;   // running max of the sweep RMS increment (Solver.cu:941, 957)
    double redSum;    // red half-sweep sum of squared increments
    double lastSum;   // last full-sweep sum
    int sweeps;       // sweeps executed in the current call
    int stop;         // relNorm < tol seen: remaining sweeps of the call are skipped
};

struct CaseState {
    unsigned counter;       // last-workgroup counter of the running reduction
    int done;               // 0 running, 1 converged, 2 diverged: every kernel returns at once
    long long iter;         // outer iterations executed
    SmoothState sm[kSlots];
    int sweepsIter[4];      // sweeps executed in the current outer iteration (u, v, w, p')
    double mass;            // sum b'^2 (calcPresCorrSrcTerm)
    double resSq[3];        // sum (b - A phi)^2 of u, v, w before the sweeps
    double upd[7];          // sums of squared increments p,u,v,w,uf,vf,wf (updateField)
    double pNormMax, uNormMax, vNormMax, wNormMax, ufNormMax, vfNormMax, wfNormMax;
    double resMax[4];
    double lastScaled;
    int lastConverged;
    int pad;
};

struct CaseParams {
    int dim, nx, ny, nz;
    long long N;
    float dx, dy, dz, Ax, Ay, Az, rho, mu;
    int bct[6];
    float bcv[6][3];
    const float *prof[6];
    float relax[4];
    double tol[4];
    double tolMass, tolRes;
    int stopRule;
    int histCap;
    float *u, *v, *w, *p, *pc, *uf, *vf, *wf;
    float *mcoef, *pcoef, *bu, *bv, *bw, *bp;
    float *tmp, *tmp2;       // p' defect and refinement correction
    float *jac;              // Jacobi destination
    double *partials;        // kPart doubles per workgroup
    CaseState *st;
    ifl_iter_record *hist;   // ring of histCap records indexed by iter % histCap
};

// ---------------- indexing (kernels.cuh:30-38, 105-114) ----------------
__device__ __forceinline__ long long cid(const CaseParams &P, int i, int j, int k) {
    return i + (long long)P.nx * (j + (long long)P.ny * k);
}
__device__ __forceinline__ long long ufi(const CaseParams &P, int i, int j, int k) {
    return i + (long long)(P.nx + 1) * (j + (long long)P.ny * k);
}
__device__ __forceinline__ long long vfi(const CaseParams &P, int i, int j, int k) {
    return i + (long long)P.nx * (j + (long long)(P.ny + 1) * k);
}
__device__ __forceinline__ long long wfi(const CaseParams &P, int i, int j, int k) {
    return i + (long long)P.nx * (j + (long long)P.ny * k);
}

// velBCs::val[f][comp] (config.h:35-42) or the per-face profile of the normal component.
__device__ __forceinline__ float bcval(const CaseParams &P, int f, int comp, long long q) {
    if (comp == (f >> 1) && P.prof[f] != nullptr) return P.prof[f][q];
    return P.bcv[f][comp];
}
// Boundary face velocity, kernels.cuh:87-93: wall 0, inlet value, outlet adjacent cell.
__device__ __forceinline__ float faceBC(const CaseParams &P, int f, long long q, float adjacent) {
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

// ---------------- FP64 reductions (wave64 butterflies, fixed order) ----------------
__device__ __forceinline__ double waveSum(double x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
    return x;
}

// Reduce K per-thread values over the workgroup; thread t < K returns the total of value t.
template <int K>
__device__ __forceinline__ double blockSum(const double (&v)[K], double *lds /* kWaves*K */) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        double w = waveSum(v[s]);
        if (lane == 0) lds[wave * K + s] = w;
    }
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x < K) {
        r = lds[threadIdx.x];
#pragma unroll
        for (int q = 1; q < kWaves; ++q) r += lds[q * K + threadIdx.x];
    }
    __syncthreads();
    return r;
}

__device__ __forceinline__ void storeAgent(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double loadAgent(const double *p) {
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publish this workgroup's K partials and return true in the last workgroup to finish.
template <int K>
__device__ __forceinline__ bool publishPartials(const CaseParams &P, const double (&v)[K], long long blk,
                                                long long nblk, double *lds) {
    double tot = blockSum<K>(v, lds);
    if (threadIdx.x < K) storeAgent(&P.partials[blk * kPart + threadIdx.x], tot);
    __shared__ int amLast;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned prev = __hip_atomic_fetch_add(&P.st->counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        amLast = (prev == (unsigned)(nblk - 1)) ? 1 : 0;
    }
    __syncthreads();
    if (amLast) __threadfence();
    return amLast != 0;
}

// In the last workgroup: total of partial s over all nblk workgroups, in a fixed order
// (thread t adds workgroups t, t+256, ... in sequence; the 256 results are combined by the
// fixed wave/LDS tree), so the result does not depend on which workgroup finished last.
template <int K>
__device__ __forceinline__ void gatherPartials(const CaseParams &P, long long nblk, double (&out)[K], double *lds) {
    double acc[K];
#pragma unroll
    for (int s = 0; s < K; ++s) acc[s] = 0.0;
    for (long long b = threadIdx.x; b < nblk; b += kBlock)
#pragma unroll
        for (int s = 0; s < K; ++s) acc[s] += loadAgent(&P.partials[b * kPart + s]);
    double tot = blockSum<K>(acc, lds);
    __shared__ double bc[K];
    if (threadIdx.x < K) bc[threadIdx.x] = tot;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < K; ++s) out[s] = bc[s];
    if (threadIdx.x == 0) __hip_atomic_store(&P.st->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int caseDone(const CaseParams &P) {
    return __hip_atomic_load(&P.st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace ifl
