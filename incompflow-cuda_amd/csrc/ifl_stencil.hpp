// ifl_stencil.hpp - per-cell coefficient formulas shared by the assembly kernels and the fused
// sweep kernels, so a coefficient recomputed inside a sweep equals the stored one bit for bit.
#pragma once

#include "ifl_device.hpp"

namespace ifl {

// The case constants of the link coefficients: density, face areas and the diffusion conductances
// mu*A/d of each direction.  A kernel reads them once (as the same float expressions, so the same
// bits); read through P after every store to global memory, the conductances' divisions would be
// recomputed per cell.
struct LinkConsts {
    float rho, Ax, Ay, Az, gx, gy, gz;
};
template <int DIM>
__device__ __forceinline__ LinkConsts linkConsts(const CaseParams &P, float mu) {
    LinkConsts k;
    k.rho = P.rho;
    k.Ax = P.Ax;
    k.Ay = P.Ay;
    k.Az = P.Az;
    k.gx = mu * P.Ax / P.dx;
    k.gy = mu * P.Ay / P.dy;
    k.gz = DIM == 3 ? mu * P.Az / P.dz : 0.0f;
    return k;
}

// Link coefficients of a transported quantity with diffusion coefficient mu from the six face
// velocities: calcMomLinkCoefKernel (kernels.cuh:116-131; T3 fixed: aT/aB in their own planes)
// with mu = P.mu for momentum and mu = P.gam (k / c_p) for temperature.
template <int DIM>
__device__ __forceinline__ void linkCoef(const LinkConsts &k, float ue, float uw, float vn, float vs, float wt,
                                         float wb, float (&a)[7]) {
    const float rho = k.rho;
    a[sE] = (rho * min0(ue) * k.Ax) - k.gx;
    a[sW] = (rho * min0(-uw) * k.Ax) - k.gx;
    a[sN] = (rho * min0(vn) * k.Ay) - k.gy;
    a[sS] = (rho * min0(-vs) * k.Ay) - k.gy;
    a[sC] = -(a[sE] + a[sW] + a[sN] + a[sS]);
    a[sT] = 0.0f;
    a[sB] = 0.0f;
    if (DIM == 3) {
        a[sT] = (rho * min0(wt) * k.Az) - k.gz;
        a[sB] = (rho * min0(-wb) * k.Az) - k.gz;
        a[sC] += -(a[sT] + a[sB]);
    }
}
template <int DIM>
__device__ __forceinline__ void linkCoef(const CaseParams &P, float mu, float ue, float uw, float vn, float vs,
                                         float wt, float wb, float (&a)[7]) {
    linkCoef<DIM>(linkConsts<DIM>(P, mu), ue, uw, vn, vs, wt, wb, a);
}

template <int DIM>
__device__ __forceinline__ void momLinkCoef(const CaseParams &P, float ue, float uw, float vn, float vs, float wt,
                                            float wb, float (&a)[7]) {
    linkCoef<DIM>(P, P.mu, ue, uw, vn, vs, wt, wb, a);
}

__device__ __forceinline__ int bcExSlot(int f) {
    return f == E ? sE : f == W ? sW : f == N ? sN : f == S ? sS : f == T ? sT : sB;
}
__device__ __forceinline__ int bcInSlot(int f) {
    return f == E ? sW : f == W ? sE : f == N ? sS : f == S ? sN : f == T ? sB : sT;
}

// Coefficient part of the boundary row of face f, applyBCsToMomEqKernel (kernels.cuh:304-331):
// a face with a prescribed value (velocity wall/inlet, temperature Dirichlet) or a zero-gradient
// face (velocity outlet, temperature Neumann).
__device__ __forceinline__ void bcRowCoefKind(bool given, int f, float (&a)[7]) {
    const int ex = bcExSlot(f), in = bcInSlot(f);
    const float aEx = a[ex];
    if (given) {                    // wall / inlet: 2nd-order one-sided wall gradient
        a[sC] -= 2 * aEx;
        a[in] += aEx / 3;
        a[ex] = 0;
    } else {                        // outlet: zero gradient
        a[sC] += aEx;
        a[ex] = 0;
    }
}

__device__ __forceinline__ void bcRowCoef(const CaseParams &P, int f, float (&a)[7]) {
    bcRowCoefKind(ghostCopy(P, f), f, a);
}
__device__ __forceinline__ bool tempDirichlet(const CaseParams &P, int f) { return P.tbct[f] == IFL_TEMP_DIRICHLET; }

// Boundary rows of the temperature equation of cell (i,j,k), in the same order.
template <int DIM>
__device__ __forceinline__ void tempBcRowsCoef(const CaseParams &P, int i, int j, int k, float (&a)[7]) {
    if (i == P.nx - 1) bcRowCoefKind(tempDirichlet(P, E), E, a);
    if (i == 0) bcRowCoefKind(tempDirichlet(P, W), W, a);
    if (j == P.ny - 1) bcRowCoefKind(tempDirichlet(P, N), N, a);
    if (j == 0) bcRowCoefKind(tempDirichlet(P, S), S, a);
    if (DIM == 3) {
        if (k == P.nz - 1) bcRowCoefKind(tempDirichlet(P, T), T, a);
        if (k == 0) bcRowCoefKind(tempDirichlet(P, B), B, a);
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
