// ifl_stencil.hpp - per-cell coefficient formulas shared by the assembly kernels and the fused
// sweep kernels, so a coefficient recomputed inside a sweep equals the stored one bit for bit.
#pragma once

#include "ifl_device.hpp"

namespace ifl {

// Momentum link coefficients from the six face velocities, calcMomLinkCoefKernel
// (kernels.cuh:116-131; T3 fixed: aT/aB in their own planes).
template <int DIM>
__device__ __forceinline__ void momLinkCoef(const CaseParams &P, float ue, float uw, float vn, float vs, float wt,
                                            float wb, float (&a)[7]) {
    const float rho = P.rho, mu = P.mu;
    a[sE] = (rho * min0(ue) * P.Ax) - mu * P.Ax / P.dx;
    a[sW] = (rho * min0(-uw) * P.Ax) - mu * P.Ax / P.dx;
    a[sN] = (rho * min0(vn) * P.Ay) - mu * P.Ay / P.dy;
    a[sS] = (rho * min0(-vs) * P.Ay) - mu * P.Ay / P.dy;
    a[sC] = -(a[sE] + a[sW] + a[sN] + a[sS]);
    a[sT] = 0.0f;
    a[sB] = 0.0f;
    if (DIM == 3) {
        a[sT] = (rho * min0(wt) * P.Az) - mu * P.Az / P.dz;
        a[sB] = (rho * min0(-wb) * P.Az) - mu * P.Az / P.dz;
        a[sC] += -(a[sT] + a[sB]);
    }
}

__device__ __forceinline__ int bcExSlot(int f) {
    return f == E ? sE : f == W ? sW : f == N ? sN : f == S ? sS : f == T ? sT : sB;
}
__device__ __forceinline__ int bcInSlot(int f) {
    return f == E ? sW : f == W ? sE : f == N ? sS : f == S ? sN : f == T ? sB : sT;
}

// Coefficient part of the boundary row of face f, applyBCsToMomEqKernel (kernels.cuh:304-331).
__device__ __forceinline__ void bcRowCoef(const CaseParams &P, int f, float (&a)[7]) {
    const int ex = bcExSlot(f), in = bcInSlot(f);
    const float aEx = a[ex];
    if (ghostCopy(P, f)) {          // wall / inlet: 2nd-order one-sided wall gradient
        a[sC] -= 2 * aEx;
        a[in] += aEx / 3;
        a[ex] = 0;
    } else {                        // outlet: zero gradient
        a[sC] += aEx;
        a[ex] = 0;
    }
}

// Boundary rows of cell (i,j,k) in launch order E, W, N, S, T, B (Solver.cu:583-608).
template <int DIM>
__device__ __forceinline__ void momBcRowsCoef(const CaseParams &P, int i, int j, int k, float (&a)[7]) {
    if (i == P.nx - 1) bcRowCoef(P, E, a);
    if (i == 0) bcRowCoef(P, W, a);
    if (j == P.ny - 1) bcRowCoef(P, N, a);
    if (j == 0) bcRowCoef(P, S, a);
    if (DIM == 3) {
        if (k == P.nz - 1) bcRowCoef(P, T, a);
        if (k == 0) bcRowCoef(P, B, a);
    }
}

}  // namespace ifl
