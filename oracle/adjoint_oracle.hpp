// adjoint_oracle.hpp - CPU reference of the discrete adjoint of the converged SIMPLE state
// (TEST INFRASTRUCTURE ONLY: loaded by the tests through oracle_capi.cpp, never by the product).
//
// The reference (IncompFlow-CUDA) has no adjoint; what it fixes is the forward iteration that
// Simple<Real> restates (Solver.cu:160-290, kernels.cuh).  Its converged state x = (u, p, uf) is
// a fixed point x = G(x, theta) of one outer iteration, where theta = (inlet normal-velocity
// profiles, mu).  For an objective J(x) (linear in u, p), the fixed-point adjoint
//     lambda = G_x^T lambda + J_x^T,   dJ/dtheta = lambda^T G_theta
// gives the gradient of J at the converged state (reverse accumulation of a fixed-point
// iteration; G_x has the forward iteration's contraction rate, so the adjoint iteration
// converges whenever the forward one did).
//
// At the fixed point, G linearises stage by stage (u* = smoothed velocity, p' = correction):
//   stage 0  boundary faces from u and theta (applyBCsToFaceVel, kernels.cuh:47-95)
//   stage 1  d(coefficients) from d(uf) and d(mu) (calcMomLinkCoef, kernels.cuh:97-134, with
//            the upwind switch min(.,0)), BC rows (kernels.cuh:252-334: the 8/3 rule), and
//            d(source) from dp (kernels.cuh:136-250)
//   stage 2  du* = n relaxed red-black sweeps on A z = db - dA u starting from du: at the fixed
//            point A u = b, so the smoother's dependence on A reduces to this residual
//   stage 3  Rhie-Chow (kernels.cuh:336-500) including the dependence of 1/aC on uf
//   stage 4  db' = div duf*, dp' = n_p relaxed sweeps from 0 (b' = 0 and p' = 0 at the fixed
//            point, so the p' matrix's own dependence drops out)
//   stage 5  correction (kernels.cuh:583-843), linear in dp' for the same reason.
// tangent() applies G_x dx + G_theta dtheta; transpose() applies (G_x^T, G_theta^T).  Both are
// pinned by dot-product tests <G dx, l> = <dx, G^T l>, tangent() by central differences of the
// oracle's own outer iteration, and the gradients by central differences of converged solves.
//
// Every transposed stage is written as a gather per cell (each output element is a sum over
// the inputs that reach it, in a fixed order), the form the device kernels use, so that the
// FP32 oracle and the device agree bit for bit at a fixed number of adjoint iterations.
#pragma once

#include "simple_oracle.hpp"

namespace oracle {

template <class Real>
struct AdjVec {
    std::vector<Real> u[3];   // cell velocity components
    std::vector<Real> p;
    std::vector<Real> f[3];   // face arrays laid out as uf, vf, wf; only interior faces are state
};

struct AdjParams {
    std::vector<double> prof[6];   // normal-velocity profile per face (inlet faces only)
    double mu = 0.0;
};

template <class Real>
class Adjoint {
public:
    Simple<Real> &s;
    int dim, nx, ny, nz;
    int64_t Nc;
    int ns;                        // coefficient slots (1 + 2 dim)
    std::vector<Real> apre;        // coefficients before the BC rows, at the fixed point
    std::vector<Real> gu[3], gp;   // objective J = sum gu_d u_d + sum gp p
    AdjVec<Real> lam;              // adjoint state
    // intermediates of the last transpose() (kept for inspection by tests)
    std::vector<Real> pcb, bpc, us[3], acb, rb[3], abar;
    AdjParams grad;                // G_theta^T lambda of the last transpose()
    double inc[7] = {0, 0, 0, 0, 0, 0, 0};   // squared increments u,v,w,p,uf,vf,wf of the last iterate()
    double incMax[7] = {0, 0, 0, 0, 0, 0, 0};
    int64_t iters = 0;

    explicit Adjoint(Simple<Real> &sim) : s(sim) {
        dim = s.dim; nx = s.nx; ny = s.ny; nz = s.nz; Nc = s.Nc;
        ns = 1 + 2 * dim;
        alloc(lam);
        for (int d = 0; d < 3; ++d) {
            gu[d].assign(Nc, 0);
            us[d].assign(Nc, 0);
            rb[d].assign(Nc, 0);
        }
        gp.assign(Nc, 0);
        pcb.assign(Nc, 0); bpc.assign(Nc, 0); acb.assign(Nc, 0);
        abar.assign(Nc * ns, 0);
    }

    void alloc(AdjVec<Real> &x) const {
        for (int d = 0; d < 3; ++d) x.u[d].assign(Nc, 0);
        x.p.assign(Nc, 0);
        x.f[0].assign(s.uf.size(), 0);
        x.f[1].assign(s.vf.size(), 0);
        x.f[2].assign(s.wf.size(), 0);
    }
    void allocParams(AdjParams &t) const {
        for (int f = 0; f < 6; ++f) t.prof[f].assign(faceCells(f), 0.0);
        t.mu = 0.0;
    }

    // ---------------- geometry helpers ----------------
    int nDir(int d) const { return d == 0 ? nx : d == 1 ? ny : nz; }
    int64_t stride(int d) const { return d == 0 ? 1 : d == 1 ? (int64_t)nx : (int64_t)nx * ny; }
    Real area(int d) const { return d == 0 ? s.areaX : d == 1 ? s.areaY : s.areaZ; }
    Real spacing(int d) const { return d == 0 ? s.dx : d == 1 ? s.dy : s.dz; }
    int idxOf(int d, int i, int j, int k) const { return d == 0 ? i : d == 1 ? j : k; }
    size_t faceCells(int f) const {
        return f < 2 ? (size_t)ny * nz : f < 4 ? (size_t)nx * nz : (size_t)nx * ny;
    }
    // position of cell (i,j,k) inside face f's profile (the q of applyBCsToFaceVel)
    int64_t faceQ(int f, int i, int j, int k) const {
        return f < 2 ? j + (int64_t)ny * k : f < 4 ? i + (int64_t)nx * k : i + (int64_t)nx * j;
    }
    // lo face (W/S/B) of cell (i,j,k) in the face array of direction d; the hi face is the lo
    // face of the next cell along d
    int64_t faceLo(int d, int i, int j, int k) const {
        return d == 0 ? s.ufi(i, j, k) : d == 1 ? s.vfi(i, j, k) : s.wfi(i, j, k);
    }
    int64_t faceHi(int d, int i, int j, int k) const {
        return d == 0 ? s.ufi(i + 1, j, k) : d == 1 ? s.vfi(i, j + 1, k) : s.wfi(i, j, k + 1);
    }
    const std::vector<Real> &faceArr(int d) const { return d == 0 ? s.uf : d == 1 ? s.vf : s.wf; }
    const std::vector<Real> &vel(int d) const { return d == 0 ? s.u : d == 1 ? s.v : s.w; }
    static int loFace(int d) { return 2 * d + 1; }
    static int hiFace(int d) { return 2 * d; }
    static int loSlot(int d) { return d == 0 ? aW : d == 1 ? aS : aB; }
    static int hiSlot(int d) { return d == 0 ? aE : d == 1 ? aN : aT; }
    Real aCf(int64_t c) const { return s.ucoef[s.slot(c, aC)]; }
    Real coefAt(int64_t c, int a) const { return s.ucoef[s.slot(c, a)]; }
    bool ghost(int f) const { return s.ghost_copy(f); }

    template <class F>
    void forCells(F &&fn) const {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) fn(i, j, k, s.cid(i, j, k));
    }

    // ---------------- the fixed point ----------------
    // Face BCs, momentum matrix (keeping the coefficients before the BC rows) and p' matrix of
    // the current state.
    void setup() {
        s.applyBCsToFaceVel();
        s.calcMomLinkCoef();
        apre = s.ucoef;
        s.calcMomSrcTerm();
        s.applyBCsToMomEq();
        s.calcPresCorrLinkCoef();
    }

    void setPressureDropObjective() {
        for (int d = 0; d < 3; ++d) std::fill(gu[d].begin(), gu[d].end(), (Real)0);
        std::fill(gp.begin(), gp.end(), (Real)0);
        const Real w = (Real)1 / (Real)((int64_t)ny * nz);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                gp[s.cid(0, j, k)] += w;
                gp[s.cid(nx - 1, j, k)] -= w;
            }
    }
    double objective() const {
        double J = 0;
        for (int64_t c = 0; c < Nc; ++c) {
            J += (double)gp[c] * (double)s.p[c];
            for (int d = 0; d < dim; ++d) J += (double)gu[d][c] * (double)vel(d)[c];
        }
        return J;
    }

    // ================= tangent: G_x dx + G_theta dtheta =================
    void tangent(const AdjVec<Real> &dx, const AdjParams &dth, AdjVec<Real> &out) {
        const int64_t N = Nc;
        // stage 0: face perturbations, boundary faces from du / dtheta
        std::vector<Real> F[3];
        for (int d = 0; d < dim; ++d) {
            F[d] = dx.f[d];
            forCells([&](int i, int j, int k, int64_t c) {
                const int idx = idxOf(d, i, j, k), n = nDir(d);
                if (idx == 0) F[d][faceLo(d, i, j, k)] = bcFacePert(loFace(d), c, d, faceQ(loFace(d), i, j, k), dx.u[d], dth);
                if (idx == n - 1) F[d][faceHi(d, i, j, k)] = bcFacePert(hiFace(d), c, d, faceQ(hiFace(d), i, j, k), dx.u[d], dth);
            });
        }
        // stage 1: coefficient and source perturbations
        std::vector<Real> dA((size_t)N * ns, 0), db[3];
        for (int d = 0; d < dim; ++d) db[d].assign(N, 0);
        const Real rho = s.density, dmu = (Real)dth.mu;
        forCells([&](int i, int j, int k, int64_t c) {
            Real sum = 0;
            for (int d = 0; d < dim; ++d) {
                const Real A = area(d), h = spacing(d);
                const Real ue = faceArr(d)[faceHi(d, i, j, k)], uw = faceArr(d)[faceLo(d, i, j, k)];
                const Real dhi = (ue < 0 ? rho * A * F[d][faceHi(d, i, j, k)] : (Real)0) - dmu * A / h;
                const Real dlo = (uw > 0 ? -(rho * A * F[d][faceLo(d, i, j, k)]) : (Real)0) - dmu * A / h;
                dA[s.slot(c, hiSlot(d))] = dhi;
                dA[s.slot(c, loSlot(d))] = dlo;
                sum += dhi + dlo;
            }
            dA[s.slot(c, aC)] = -sum;
            for (int d = 0; d < dim; ++d) db[d][c] = presSrcPert(d, i, j, k, dx.p);
        });
        // BC rows, faces in launch order E, W, N, S, T, B
        static const int exSlot[6] = {aE, aW, aN, aS, aT, aB};
        static const int inSlot[6] = {aW, aE, aS, aN, aB, aT};
        for (int f = 0; f < 2 * dim; ++f)
            forBoundary(f, [&](int i, int j, int k, int64_t c, int64_t q) {
                const int64_t iC = s.slot(c, aC), iEx = s.slot(c, exSlot[f]), iIn = s.slot(c, inSlot[f]);
                if (s.bct[f] == wall || s.bct[f] == inlet) {
                    const Real aEx = apre[iEx];
                    for (int comp = 0; comp < dim; ++comp) {
                        const Real phib = s.bcval(f, comp, q);
                        const Real dphib = (s.bct[f] == inlet && comp == normal_dir(f)) ? (Real)dth.prof[f][q] : (Real)0;
                        db[comp][c] -= (Real)8 / 3 * (dA[iEx] * phib + aEx * dphib);
                    }
                    dA[iC] -= 2 * dA[iEx];
                    dA[iIn] += dA[iEx] / 3;
                    dA[iEx] = 0;
                } else {
                    dA[iC] += dA[iEx];
                    dA[iEx] = 0;
                }
                (void)i; (void)j; (void)k;
            });
        // stage 2: momentum sweeps on A z = db - dA u from du
        std::vector<Real> ustar[3];
        for (int d = 0; d < dim; ++d) {
            std::vector<Real> r(N);
            forCells([&](int i, int j, int k, int64_t c) { r[c] = db[d][c] - applyPert(dA, vel(d), i, j, k, c); });
            ustar[d] = dx.u[d];
            for (int it = 0; it < s.nIter[d]; ++it) s.sweep(ustar[d].data(), s.ucoef.data(), r.data(), s.relax[d], IFL_SMOOTHER_RBGS);
        }
        // stage 3: Rhie-Chow (interior faces) and the face BCs from du*
        std::vector<Real> Fs[3];
        for (int d = 0; d < dim; ++d) {
            Fs[d].assign(faceArr(d).size(), 0);
            forCells([&](int i, int j, int k, int64_t c) {
                const int idx = idxOf(d, i, j, k), n = nDir(d);
                if (idx >= 1) Fs[d][faceLo(d, i, j, k)] = rcPert(d, c, idx, ustar[d], dx.p, dA);
                if (idx == 0) Fs[d][faceLo(d, i, j, k)] = bcFacePert(loFace(d), c, d, faceQ(loFace(d), i, j, k), ustar[d], dth);
                if (idx == n - 1) Fs[d][faceHi(d, i, j, k)] = bcFacePert(hiFace(d), c, d, faceQ(hiFace(d), i, j, k), ustar[d], dth);
            });
        }
        // stage 4: db' = div duf*, dp' from n_p sweeps
        std::vector<Real> dbp(N), dpc(N, 0);
        forCells([&](int i, int j, int k, int64_t c) {
            Real b = 0;
            for (int d = 0; d < dim; ++d)
                b += s.density * area(d) * (Fs[d][faceLo(d, i, j, k)] - Fs[d][faceHi(d, i, j, k)]);
            dbp[c] = b;
        });
        for (int it = 0; it < s.nIter[3]; ++it) s.sweep(dpc.data(), s.pcoef.data(), dbp.data(), s.relax[3], IFL_SMOOTHER_RBGS);
        // stage 5: corrections
        alloc(out);
        forCells([&](int i, int j, int k, int64_t c) {
            out.p[c] = dx.p[c] + s.relax[3] * dpc[c];
            for (int d = 0; d < dim; ++d) {
                const int idx = idxOf(d, i, j, k), n = nDir(d);
                const int64_t st = stride(d);
                Real L, R;
                if (idx == 0) { L = ghost(loFace(d)) ? dpc[c] : (Real)0; R = dpc[c + st]; }
                else if (idx == n - 1) { R = ghost(hiFace(d)) ? dpc[c] : (Real)0; L = dpc[c - st]; }
                else { L = dpc[c - st]; R = dpc[c + st]; }
                out.u[d][c] = ustar[d][c] + s.relax[d] * (Real)0.5 * area(d) / aCf(c) * (L - R);
                if (idx >= 1) {
                    const int64_t Lc = c - st;
                    out.f[d][faceLo(d, i, j, k)] = Fs[d][faceLo(d, i, j, k)]
                        + s.relax[d] * (Real)0.5 * (1 / aCf(Lc) + 1 / aCf(c)) * area(d) * (dpc[Lc] - dpc[c]);
                }
            }
        });
    }

    // ================= transpose: (G_x^T lambda, G_theta^T lambda) =================
    void transpose(const AdjVec<Real> &l, AdjVec<Real> &out, AdjParams &tb) {
        const int64_t N = Nc;
        allocParams(tb);
        // T5: p'-bar from the corrections (gather per cell)
        forCells([&](int i, int j, int k, int64_t c) { pcb[c] = gatherPcorr(l, i, j, k, c); });
        // T4: transposed p' sweeps from (p'-bar, 0); b'-bar accumulates
        std::vector<Real> ph = pcb;
        std::fill(bpc.begin(), bpc.end(), (Real)0);
        for (int it = 0; it < s.nIter[3]; ++it) sweepT(ph, bpc, s.pcoef.data(), s.relax[3]);
        // T3 (and the divergence part of T4): u*-bar, p-bar partial, aC-bar, inlet theta-bar
        std::vector<Real> pacc(N);
        forCells([&](int i, int j, int k, int64_t c) {
            for (int d = 0; d < dim; ++d) us[d][c] = gatherUstar(l, d, i, j, k, c);
            pacc[c] = gatherRcP(l, i, j, k, c);
            acb[c] = gatherRcAc(l, i, j, k, c);
            for (int d = 0; d < dim; ++d) {
                const int idx = idxOf(d, i, j, k), n = nDir(d);
                const Real rA = s.density * area(d);
                if (idx == 0 && s.bct[loFace(d)] == inlet) tb.prof[loFace(d)][faceQ(loFace(d), i, j, k)] += (double)(rA * bpc[c]);
                if (idx == n - 1 && s.bct[hiFace(d)] == inlet)
                    tb.prof[hiFace(d)][faceQ(hiFace(d), i, j, k)] += (double)(-(rA * bpc[c]));
            }
        });
        // T2: transposed momentum sweeps from (u*-bar, 0)
        alloc(out);
        for (int d = 0; d < dim; ++d) {
            std::vector<Real> phd = us[d];
            std::fill(rb[d].begin(), rb[d].end(), (Real)0);
            for (int it = 0; it < s.nIter[d]; ++it) sweepT(phd, rb[d], s.ucoef.data(), s.relax[d]);
            out.u[d] = phd;
        }
        // T1: coefficient adjoints (before the BC rows), p-bar from the pressure sources, objective
        double mub = 0;
        forCells([&](int i, int j, int k, int64_t c) {
            coefBar(i, j, k, c, &abar[(size_t)c * ns], tb);
            out.p[c] = pacc[c] + gatherPresSrc(i, j, k, c) + gp[c];
            mub += muPart(&abar[(size_t)c * ns]);
        });
        tb.mu = mub;
        // T0/T1: face adjoints from the coefficients; boundary faces to u (outlet) or theta (inlet)
        forCells([&](int i, int j, int k, int64_t c) {
            for (int d = 0; d < dim; ++d) {
                const int idx = idxOf(d, i, j, k), n = nDir(d);
                Real o = out.u[d][c];
                if (idx == 0) {
                    const Real g = faceCoefBar(d, -1, c, faceArr(d)[faceLo(d, i, j, k)]);
                    if (s.bct[loFace(d)] == outlet) o += g;
                    else if (s.bct[loFace(d)] == inlet) tb.prof[loFace(d)][faceQ(loFace(d), i, j, k)] += (double)g;
                }
                if (idx == n - 1) {
                    const Real g = faceCoefBar(d, +1, c, faceArr(d)[faceHi(d, i, j, k)]);
                    if (s.bct[hiFace(d)] == outlet) o += g;
                    else if (s.bct[hiFace(d)] == inlet) tb.prof[hiFace(d)][faceQ(hiFace(d), i, j, k)] += (double)g;
                }
                out.u[d][c] = o + gu[d][c];
                if (idx >= 1) out.f[d][faceLo(d, i, j, k)] = faceBarInterior(d, c - stride(d), c, faceArr(d)[faceLo(d, i, j, k)]);
            }
        });
    }

    // One adjoint iteration lambda <- G_x^T lambda + J_x^T; returns the scaled increment.
    double iterate() {
        AdjVec<Real> nl;
        transpose(lam, nl, grad);
        for (int q = 0; q < 7; ++q) inc[q] = 0;
        for (int64_t c = 0; c < Nc; ++c) {
            for (int d = 0; d < dim; ++d) {
                const double e = (double)nl.u[d][c] - (double)lam.u[d][c];
                inc[d] += e * e;
            }
            const double e = (double)nl.p[c] - (double)lam.p[c];
            inc[3] += e * e;
        }
        for (int d = 0; d < dim; ++d)
            for (size_t q = 0; q < nl.f[d].size(); ++q) {
                const double e = (double)nl.f[d][q] - (double)lam.f[d][q];
                inc[4 + d] += e * e;
            }
        lam = nl;
        double sc = 0;
        for (int q = 0; q < 7; ++q) {
            if (iters < 5) incMax[q] = std::max(incMax[q], inc[q]);
            if (incMax[q] > 0) sc = std::max(sc, std::sqrt(inc[q] / incMax[q]));
        }
        ++iters;
        return sc;
    }

    // ---------------- building blocks ----------------
    template <class F>
    void forBoundary(int f, F &&fn) const {
        if (f == E || f == W) {
            const int i = f == E ? nx - 1 : 0;
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j) fn(i, j, k, s.cid(i, j, k), j + (int64_t)ny * k);
        } else if (f == N || f == S) {
            const int j = f == N ? ny - 1 : 0;
            for (int k = 0; k < nz; ++k)
                for (int i = 0; i < nx; ++i) fn(i, j, k, s.cid(i, j, k), i + (int64_t)nx * k);
        } else {
            const int k = f == T ? nz - 1 : 0;
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) fn(i, j, k, s.cid(i, j, k), i + (int64_t)nx * j);
        }
    }

    // boundary face perturbation (faceBC linearised): wall 0, inlet d(theta), outlet d(adjacent)
    Real bcFacePert(int f, int64_t c, int d, int64_t q, const std::vector<Real> &du, const AdjParams &dth) const {
        (void)d;
        if (s.bct[f] == wall) return 0;
        if (s.bct[f] == inlet) return dth.prof[f].empty() ? (Real)0 : (Real)dth.prof[f][q];
        return du[c];
    }

    Real presSrcPert(int d, int i, int j, int k, const std::vector<Real> &dp) const {
        const int idx = idxOf(d, i, j, k), n = nDir(d);
        const int64_t st = stride(d), c = s.cid(i, j, k);
        Real L, R;
        if (idx == 0) { L = ghost(loFace(d)) ? dp[c] : (Real)0; R = dp[c + st]; }
        else if (idx == n - 1) { R = ghost(hiFace(d)) ? dp[c] : (Real)0; L = dp[c - st]; }
        else { L = dp[c - st]; R = dp[c + st]; }
        return (Real)0.5 * (L - R) * area(d);
    }

    // (dA phi)_c with out-of-grid neighbours 0
    Real applyPert(const std::vector<Real> &dA, const std::vector<Real> &phi, int i, int j, int k, int64_t c) const {
        Real r = dA[s.slot(c, aC)] * phi[c];
        for (int d = 0; d < dim; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(d);
            const int64_t st = stride(d);
            if (idx != n - 1) r += dA[s.slot(c, hiSlot(d))] * phi[c + st];
            if (idx != 0) r += dA[s.slot(c, loSlot(d))] * phi[c - st];
        }
        return r;
    }

    // pressure values of the Rhie-Chow stencil of the face between c-st and c (idx = c's index)
    void rcStencil(int d, int64_t c, int idx, const std::vector<Real> &p, Real &pL, Real &pR, Real &pLL, Real &pRR) const {
        const int n = nDir(d);
        const int64_t st = stride(d), L = c - st;
        pL = p[L];
        pR = p[c];
        if (idx == 1) { pLL = ghost(loFace(d)) ? pL : (Real)0; pRR = p[c + st]; }
        else if (idx == n - 1) { pRR = ghost(hiFace(d)) ? pR : (Real)0; pLL = p[L - st]; }
        else { pLL = p[L - st]; pRR = p[c + st]; }
    }

    Real rcPert(int d, int64_t c, int idx, const std::vector<Real> &du, const std::vector<Real> &dp,
                const std::vector<Real> &dA) const {
        const int64_t st = stride(d), L = c - st;
        const Real A = area(d);
        const Real aL = aCf(L), aR = aCf(c);
        const Real daL = dA[s.slot(L, aC)], daR = dA[s.slot(c, aC)];
        Real pL, pR, pLL, pRR, dL, dR, dLL, dRR;
        rcStencil(d, c, idx, s.p, pL, pR, pLL, pRR);
        rcStencil(d, c, idx, dp, dL, dR, dLL, dRR);
        Real r = (Real)0.5 * (du[L] + du[c]);
        r += (dR - dLL) * A / (4 * aL) - (pR - pLL) * A * daL / (4 * aL * aL);
        r += (dRR - dL) * A / (4 * aR) - (pRR - pL) * A * daR / (4 * aR * aR);
        r += (Real)0.5 * (1 / aL + 1 / aR) * (dL - dR) * A;
        r -= (Real)0.5 * (daL / (aL * aL) + daR / (aR * aR)) * (pL - pR) * A;
        return r;
    }

    // One transposed red-black sweep (black half transposed, then red half transposed).
    // Half sweep of colour col: t = relax*phb, S = t/aC on col cells; other cells subtract the
    // neighbours' S times the neighbour's coefficient towards them; col cells add S to the
    // source adjoint and keep phb - t.
    void sweepT(std::vector<Real> &phb, std::vector<Real> &bb, const Real *cf, Real rlx) const {
        std::vector<Real> S(Nc);
        for (int col = 1; col >= 0; --col) {
            forCells([&](int i, int j, int k, int64_t c) {
                if (((i + j + k) & 1) == col) S[c] = rlx * phb[c] / cf[s.slot(c, aC)];
            });
            std::vector<Real> nw(phb);
            forCells([&](int i, int j, int k, int64_t c) {
                if (((i + j + k) & 1) == col) {
                    bb[c] += S[c];
                    nw[c] = phb[c] - rlx * phb[c];
                } else {
                    Real acc = phb[c];
                    for (int d = 0; d < dim; ++d) {
                        const int idx = idxOf(d, i, j, k), n = nDir(d);
                        const int64_t st = stride(d);
                        if (idx != n - 1) acc -= cf[s.slot(c + st, loSlot(d))] * S[c + st];
                        if (idx != 0) acc -= cf[s.slot(c - st, hiSlot(d))] * S[c - st];
                    }
                    nw[c] = acc;
                }
            });
            phb.swap(nw);
        }
    }

    // T5 gather: p'-bar at cell c
    Real gatherPcorr(const AdjVec<Real> &l, int i, int j, int k, int64_t c) const {
        Real acc = s.relax[3] * l.p[c];
        for (int d = 0; d < dim; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(d);
            const int64_t st = stride(d);
            const Real A = area(d), rl = s.relax[d];
            auto gcell = [&](int64_t e) { return rl * (Real)0.5 * A / aCf(e) * l.u[d][e]; };
            if (idx <= n - 2) acc += gcell(c + st);                 // c is L of cell c+st
            if (idx >= 1) acc -= gcell(c - st);                     // c is R of cell c-st
            if (idx == 0 && ghost(loFace(d))) acc += gcell(c);      // ghost L of c itself
            if (idx == n - 1 && ghost(hiFace(d))) acc -= gcell(c);  // ghost R of c itself
            auto hface = [&](int64_t Lc, int64_t Rc, int64_t fi) {
                return rl * (Real)0.5 * (1 / aCf(Lc) + 1 / aCf(Rc)) * A * l.f[d][fi];
            };
            if (idx <= n - 2) acc += hface(c, c + st, faceHi(d, i, j, k));   // c is L of its hi face
            if (idx >= 1) acc -= hface(c - st, c, faceLo(d, i, j, k));       // c is R of its lo face
        }
        return acc;
    }

    // u*-bar of component d at face fi between cells: lambda face + divergence adjoint
    Real faceStarBar(const AdjVec<Real> &l, int d, int64_t Lc, int64_t Rc, int64_t fi) const {
        return l.f[d][fi] + s.density * area(d) * (bpc[Rc] - bpc[Lc]);
    }

    // T3 gather: u*-bar_d at cell c
    Real gatherUstar(const AdjVec<Real> &l, int d, int i, int j, int k, int64_t c) const {
        const int idx = idxOf(d, i, j, k), n = nDir(d);
        const int64_t st = stride(d);
        Real acc = l.u[d][c];
        if (idx >= 1) acc += (Real)0.5 * faceStarBar(l, d, c - st, c, faceLo(d, i, j, k));
        if (idx <= n - 2) acc += (Real)0.5 * faceStarBar(l, d, c, c + st, faceHi(d, i, j, k));
        const Real rA = s.density * area(d);
        if (idx == 0 && s.bct[loFace(d)] == outlet) acc += rA * bpc[c];
        if (idx == n - 1 && s.bct[hiFace(d)] == outlet) acc += -(rA * bpc[c]);
        return acc;
    }

    // Rhie-Chow weights of the face with R-cell index fidx (1..n-1) along d, R-cell id Rc
    void rcWeights(int d, int64_t Rc, Real &kL, Real &kR, Real &m) const {
        const int64_t Lc = Rc - stride(d);
        const Real A = area(d);
        kL = A / (4 * aCf(Lc));
        kR = A / (4 * aCf(Rc));
        m = (Real)0.5 * (1 / aCf(Lc) + 1 / aCf(Rc)) * A;
    }

    // T3 gather: p-bar at cell c (lambda_p + Rhie-Chow pressure terms)
    Real gatherRcP(const AdjVec<Real> &l, int i, int j, int k, int64_t c) const {
        Real acc = l.p[c];
        for (int d = 0; d < dim; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(d);
            const int64_t st = stride(d);
            Real kL, kR, m;
            auto w = [&](int64_t Rc) { return faceStarBar(l, d, Rc - st, Rc, faceLoOf(d, Rc)); };
            if (idx >= 1) {                       // R of face idx
                rcWeights(d, c, kL, kR, m);
                acc += (kL - m) * w(c);
            }
            if (idx <= n - 2) {                   // L of face idx+1
                rcWeights(d, c + st, kL, kR, m);
                acc += (m - kR) * w(c + st);
            }
            if (idx + 2 <= n - 1) {               // LL of face idx+2
                rcWeights(d, c + 2 * st, kL, kR, m);
                acc -= kL * w(c + 2 * st);
            }
            if (idx - 1 >= 1) {                   // RR of face idx-1
                rcWeights(d, c - st, kL, kR, m);
                acc += kR * w(c - st);
            }
            if (idx == 0 && ghost(loFace(d))) {   // ghost LL of face 1 is cell 0
                rcWeights(d, c + st, kL, kR, m);
                acc -= kL * w(c + st);
            }
            if (idx == n - 1 && ghost(hiFace(d))) {   // ghost RR of face n-1 is cell n-1
                rcWeights(d, c, kL, kR, m);
                acc += kR * w(c);
            }
        }
        return acc;
    }

    int64_t faceLoOf(int d, int64_t c) const {
        const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((int64_t)nx * ny));
        return faceLo(d, i, j, k);
    }

    // T3 gather: aC-bar at cell c (dependence of the Rhie-Chow weights on aC)
    Real gatherRcAc(const AdjVec<Real> &l, int i, int j, int k, int64_t c) const {
        Real acc = 0;
        for (int d = 0; d < dim; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(d);
            const int64_t st = stride(d);
            const Real A = area(d), a = aCf(c);
            Real pL, pR, pLL, pRR;
            if (idx <= n - 2) {                   // c is L of face idx+1
                rcStencil(d, c + st, idx + 1, s.p, pL, pR, pLL, pRR);
                const Real w = faceStarBar(l, d, c, c + st, faceHi(d, i, j, k));
                acc += w * (-((pR - pLL) * A / (4 * a * a)) - (Real)0.5 * (pL - pR) * A / (a * a));
            }
            if (idx >= 1) {                       // c is R of face idx
                rcStencil(d, c, idx, s.p, pL, pR, pLL, pRR);
                const Real w = faceStarBar(l, d, c - st, c, faceLo(d, i, j, k));
                acc += w * (-((pRR - pL) * A / (4 * a * a)) - (Real)0.5 * (pL - pR) * A / (a * a));
            }
        }
        return acc;
    }

    // T1 gather: p-bar at cell c from the pressure sources of the momentum equations
    Real gatherPresSrc(int i, int j, int k, int64_t c) const {
        Real acc = 0;
        for (int d = 0; d < dim; ++d) {
            const int idx = idxOf(d, i, j, k), n = nDir(d);
            const int64_t st = stride(d);
            const Real A = area(d);
            if (idx <= n - 2) acc += (Real)0.5 * A * rb[d][c + st];                     // L of c+st
            if (idx >= 1) acc -= (Real)0.5 * A * rb[d][c - st];                         // R of c-st
            if (idx == 0 && ghost(loFace(d))) acc += (Real)0.5 * A * rb[d][c];
            if (idx == n - 1 && ghost(hiFace(d))) acc -= (Real)0.5 * A * rb[d][c];
        }
        return acc;
    }

    // T1: adjoint of the coefficients before the BC rows at cell c (ab[0..ns)), plus the inlet
    // theta-bar of the 8/3 source term
    void coefBar(int i, int j, int k, int64_t c, Real *ab, AdjParams &tb) const {
        // adjoint of the final coefficients: -rbar_d (c) * u_d(neighbour); aC also gets aC-bar
        for (int a = 0; a < ns; ++a) ab[a] = 0;
        for (int d = 0; d < dim; ++d) {
            const std::vector<Real> &ud = vel(d);
            ab[aC] -= rb[d][c] * ud[c];
            for (int e = 0; e < dim; ++e) {
                const int idx = idxOf(e, i, j, k), n = nDir(e);
                const int64_t st = stride(e);
                if (idx != n - 1) ab[hiSlot(e)] -= rb[d][c] * ud[c + st];
                if (idx != 0) ab[loSlot(e)] -= rb[d][c] * ud[c - st];
            }
        }
        ab[aC] += acb[c];
        // BC rows transposed: the exterior slot's adjoint follows from the slots it fed
        static const int exSlot[6] = {aE, aW, aN, aS, aT, aB};
        static const int inSlot[6] = {aW, aE, aS, aN, aB, aT};
        for (int f = 0; f < 2 * dim; ++f) {
            const int d = normal_dir(f);
            const int idx = idxOf(d, i, j, k), n = nDir(d);
            const bool on = (f == hiFace(d)) ? idx == n - 1 : idx == 0;
            if (!on) continue;
            const int64_t q = faceQ(f, i, j, k);
            if (s.bct[f] == wall || s.bct[f] == inlet) {
                Real a = -2 * ab[aC] + ab[inSlot[f]] / 3;
                for (int comp = 0; comp < dim; ++comp) a -= (Real)8 / 3 * s.bcval(f, comp, q) * rb[comp][c];
                ab[exSlot[f]] = a;
                if (s.bct[f] == inlet)
                    tb.prof[f][q] += (double)(-((Real)8 / 3 * apre[s.slot(c, exSlot[f])] * rb[normal_dir(f)][c]));
            } else {
                ab[exSlot[f]] = ab[aC];
            }
        }
    }

    double muPart(const Real *ab) const {
        double m = 0;
        for (int d = 0; d < dim; ++d) {
            const double w = (double)area(d) / (double)spacing(d);
            m += w * ((double)ab[aC] - (double)ab[hiSlot(d)]);
            m += w * ((double)ab[aC] - (double)ab[loSlot(d)]);
        }
        return m;
    }

    // adjoint of a face velocity through the coefficients of one cell: side -1 = the cell's lo
    // face (it feeds the lo slot), +1 = its hi face
    Real faceCoefBar(int d, int side, int64_t c, Real ufv) const {
        const Real *ab = &abar[(size_t)c * ns];
        const Real rA = s.density * area(d);
        if (side < 0) return ufv > 0 ? rA * (ab[aC] - ab[loSlot(d)]) : (Real)0;
        return ufv < 0 ? rA * (ab[hiSlot(d)] - ab[aC]) : (Real)0;
    }
    Real faceBarInterior(int d, int64_t Lc, int64_t Rc, Real ufv) const {
        const Real *aL = &abar[(size_t)Lc * ns], *aR = &abar[(size_t)Rc * ns];
        Real a = 0;
        if (ufv < 0) a += aL[hiSlot(d)] - aL[aC];
        if (ufv > 0) a += aR[aC] - aR[loSlot(d)];
        return s.density * area(d) * a;
    }
};

}  // namespace oracle
