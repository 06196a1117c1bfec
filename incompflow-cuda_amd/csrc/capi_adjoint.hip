// capi_adjoint.hip - host side of the discrete adjoint (ifl_adjoint, ifl_adjoint_lambda).
#include "capi_host.hpp"

using namespace ifl_host;

// ---------------- discrete adjoint ----------------
namespace {

int adjAlloc(ifl_solver *h) {
    if (h->adj) return IFL_OK;
    AdjHost *a = new (std::nothrow) AdjHost();
    if (!a) return fail(IFL_ERR_NOMEM, "out of host memory");
    h->adj = a;
    const ifl::CaseParams &P = h->P;
    const size_t N = (size_t)P.N, nx = P.nx, ny = P.ny, nz = P.nz;
    const size_t nblk = (size_t)ifl::adj_blocks(P);
    ifl::AdjBufs &B = a->B;
    struct Part { void **dst; size_t bytes; };
    std::vector<Part> parts;
    for (int d = 0; d < 3; ++d) {
        parts.push_back({(void **)&B.lu[d], N * 4});
        parts.push_back({(void **)&B.us[d], N * 4});
        parts.push_back({(void **)&B.usAlt[d], N * 4});
        parts.push_back({(void **)&B.rb[d], N * 4});
    }
    parts.push_back({(void **)&B.lf[0], (nx + 1) * ny * nz * 4});
    parts.push_back({(void **)&B.lf[1], nx * (ny + 1) * nz * 4});
    parts.push_back({(void **)&B.lf[2], nx * ny * (nz + 1) * 4});
    parts.push_back({(void **)&B.lp, N * 4});
    parts.push_back({(void **)&B.pcb, N * 4});
    parts.push_back({(void **)&B.pcbAlt, N * 4});
    parts.push_back({(void **)&B.bpc, N * 4});
    parts.push_back({(void **)&B.acb, N * 4});
    parts.push_back({(void **)&B.pacc, N * 4});
    parts.push_back({(void **)&B.abar, N * 7 * 4});
    for (int q = 0; q < 4; ++q) parts.push_back({(void **)&a->gw[q], N * 4});
    const size_t faceCells[6] = {ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny};
    for (int f = 0; f < 6; ++f) parts.push_back({(void **)&B.gprof[f], faceCells[f] * 8});
    parts.push_back({(void **)&B.part, nblk * 8 * 8});
    parts.push_back({(void **)&B.st, sizeof(ifl::AdjState)});
    size_t total = 0;
    for (auto &p : parts) total += align256(p.bytes);
    IFL_HIP(hipMalloc((void **)&a->arena, total));
    size_t off = 0;
    for (auto &p : parts) {
        *p.dst = a->arena + off;
        off += align256(p.bytes);
    }
    IFL_HIP(hipMemsetAsync(a->arena, 0, total, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    B.nblk = (int)nblk;
    return IFL_OK;
}

// Tile of the 2D transposed sweeps: the configured one, else 32x16 with four sweeps per launch
// (the fastest on the 1024x256 channel once the adjoint cell kernels run one row per thread:
// 0.155 ms per adjoint iteration against 0.170-0.196 ms for the other shapes), or 32x8 with four
// when the grid has fewer than 256 such tiles.
int adjShape(const ifl::CaseParams &P, int tile) {
    if (tile > 0) return tile - 1;
    const long tiles = (long)((P.nx + 31) / 32) * ((P.ny + 15) / 16);
    return tiles >= 256 ? 1 : 3;
}

// n transposed sweeps of one equation from *x (the result ends in *x; *alt is its partner).
void enqueueAdjSweeps(const ifl::Lc &L, const ifl::AdjBufs &B, int tile, float *&x, float *&alt, float *bb,
                      const float *coef, float rlx, int n) {
    const ifl::CaseParams &P = *L.P;
    if (P.dim == 2) {
        const int shape = adjShape(P, tile), K = ifl::adj_tile_sweeps(shape);
        for (int s = 0; s < n; s += K) {
            ifl::launch_adj_tile2d(L, shape, x, alt, bb, coef, rlx, std::min(K, n - s));
            std::swap(x, alt);
        }
        return;
    }
    for (int s = 0; s < n; ++s) {
        ifl::launch_adj_sweep_t(L, B, x, alt, bb, coef, rlx, 1);
        ifl::launch_adj_sweep_t(L, B, alt, x, bb, coef, rlx, 0);
    }
}

// One adjoint iteration (reverse stage order of one outer iteration).  B is a copy: the sweeps
// swap its ping-pong pointers, and every iteration starts from the same assignment.
void enqueueAdjoint(ifl_solver *h, int tile) {
    const Plan pl = planOf(h);
    const ifl::Lc &L = pl.L;
    ifl::AdjBufs B = h->adj->B;
    const ifl::CaseParams &P = h->P;
    ifl::launch_adj_pcorr(L, B);
    enqueueAdjSweeps(L, B, tile, B.pcb, B.pcbAlt, B.bpc, P.pcoef, P.relax[3], h->cfg.n_iter[3]);
    ifl::launch_adj_rc(L, B);
    for (int d = 0; d < P.dim; ++d)
        enqueueAdjSweeps(L, B, tile, B.us[d], B.usAlt[d], B.rb[d], P.mcoef, P.relax[d], h->cfg.n_iter[d]);
    ifl::launch_adj_coef(L, B);
    ifl::launch_adj_face(L, B);
    ifl::launch_adj_reduce(L, B);
}

}  // namespace

extern "C" void ifl_adjoint_opts_default(ifl_adjoint_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->objective = IFL_OBJ_PRESSURE_DROP;
    o->max_iter = 200000;
    o->tol = 1e-7;
    o->check_every = 64;
}

extern "C" int ifl_adjoint(ifl_solver *h, const ifl_adjoint_opts *o, int face, double *dJ_dprofile, size_t n_profile,
                           double *dJ_dmu, ifl_adjoint_report *report) {
    if (!h || !o) return fail(IFL_ERR_ARG, "null argument");
    if (o->objective != IFL_OBJ_PRESSURE_DROP && o->objective != IFL_OBJ_LINEAR) return fail(IFL_ERR_ARG, "bad objective");
    if (o->tile < 0 || o->tile > ifl::kAdjShapes) return fail(IFL_ERR_ARG, "adjoint tile must be 0..6");
    const ifl_config &c = h->cfg;
    const size_t N = (size_t)h->P.N;
    if (dJ_dprofile) {
        if (face < 0 || face >= 2 * c.dim || c.bc_type[face] != IFL_INLET)
            return fail(IFL_ERR_ARG, "the profile gradient needs an inlet face");
        const size_t need = face < 2 ? (size_t)c.ny * c.nz : face < 4 ? (size_t)c.nx * c.nz : (size_t)c.nx * c.ny;
        if (n_profile != need) return fail(IFL_ERR_ARG, "profile gradient length must equal the cells on the face");
    }
    if (report) std::memset(report, 0, sizeof(*report));
    int st = adjAlloc(h);
    if (st != IFL_OK) return st;
    AdjHost &A = *h->adj;
    ifl::AdjBufs &B = A.B;
    // the fixed point: face BCs, momentum and p' matrices of the current state
    if ((st = armCase(h, h->stream, 0)) != IFL_OK) return st;
    const Plan pl = planOf(h);
    ifl::launch_face_bc(pl.L);
    ifl::launch_mom_assemble(pl.L);
    ifl::launch_pcorr_assemble(pl.L);
    IFL_HIP(hipGetLastError());
    // objective weights
    std::vector<float> w(N);
    for (int q = 0; q < 4; ++q) {
        const float *src = nullptr;
        if (o->objective == IFL_OBJ_LINEAR) src = o->weights[q];
        else if (q == 3) {
            std::fill(w.begin(), w.end(), 0.0f);
            const float v = 1.0f / (float)((long long)c.ny * c.nz);
            for (int k = 0; k < c.nz; ++k)
                for (int j = 0; j < c.ny; ++j) {
                    w[(size_t)c.nx * (j + (size_t)c.ny * k)] += v;
                    w[(size_t)c.nx - 1 + (size_t)c.nx * (j + (size_t)c.ny * k)] -= v;
                }
            src = w.data();
        }
        if (src) {
            IFL_HIP(hipMemcpyAsync(A.gw[q], src, N * 4, hipMemcpyHostToDevice, h->stream));
            IFL_HIP(hipStreamSynchronize(h->stream));
        }
        if (q < 3) B.gu[q] = src ? A.gw[q] : nullptr;
        else B.gp = src ? A.gw[q] : nullptr;
    }
    if (!o->warm_start) {
        const size_t nl = (size_t)((char *)B.pcb - (char *)B.lf[0]);   // lf[0..2] and lp are contiguous
        IFL_HIP(hipMemsetAsync(B.lf[0], 0, nl, h->stream));
        for (int d = 0; d < 3; ++d) IFL_HIP(hipMemsetAsync(B.lu[d], 0, N * 4, h->stream));
        IFL_HIP(hipMemsetAsync(B.st, 0, sizeof(ifl::AdjState), h->stream));
        IFL_HIP(hipStreamSynchronize(h->stream));
    }
    // graph of one adjoint iteration (captured per call: sweep counts and weight pointers are
    // part of it)
    if (A.graph) {
        (void)hipGraphExecDestroy(A.graph);
        A.graph = nullptr;
    }
    hipGraph_t g = nullptr;
    IFL_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueueAdjoint(h, o->tile);
    IFL_HIP(hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(&A.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(IFL_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    const int maxIter = o->max_iter > 0 ? o->max_iter : 200000;
    const double tol = o->tol > 0 ? o->tol : (o->tol < 0 ? 0.0 : 1e-7);
    const int chunk = o->check_every > 0 ? o->check_every : 64;
    IFL_HIP(hipEventRecord(h->ev0, h->stream));
    ifl::AdjState as{};
    long long done = 0;
    bool conv = false;
    double prevMub = 0.0, change = -1.0;
    // the grad_tol rule watches dJ/dmu and, when the call asks for it, the profile gradient
    std::vector<double> gNow(dJ_dprofile ? n_profile : 0), gPrev(gNow.size(), 0.0);
    while (done < maxIter) {
        const int n = (int)std::min<long long>(chunk, maxIter - done);
        for (int q = 0; q < n; ++q) IFL_HIP(hipGraphLaunch(A.graph, h->stream));
        done += n;
        IFL_HIP(hipMemcpyAsync(&as, B.st, sizeof(as), hipMemcpyDeviceToHost, h->stream));
        IFL_HIP(hipStreamSynchronize(h->stream));
        if (!std::isfinite(as.scaled) || !std::isfinite(as.mub)) return fail(IFL_ERR_DIVERGED, "non-finite adjoint");
        if (as.iters >= 5 && as.scaled < tol) {
            conv = true;
            break;
        }
        if (o->grad_tol > 0) {
            double dg = std::fabs(as.mub - prevMub), ng = std::fabs(as.mub);
            bool settled = dg <= o->grad_tol * ng;
            if (!gNow.empty()) {
                IFL_HIP(hipMemcpyAsync(gNow.data(), B.gprof[face], n_profile * 8, hipMemcpyDeviceToHost, h->stream));
                IFL_HIP(hipStreamSynchronize(h->stream));
                double d2 = 0.0, g2 = 0.0;
                for (size_t q = 0; q < n_profile; ++q) {
                    d2 += (gNow[q] - gPrev[q]) * (gNow[q] - gPrev[q]);
                    g2 += gNow[q] * gNow[q];
                }
                settled = settled && std::sqrt(d2) <= o->grad_tol * std::sqrt(g2);
                dg = std::max(dg / (ng > 0 ? ng : 1.0), std::sqrt(d2) / (g2 > 0 ? std::sqrt(g2) : 1.0));
                gPrev.swap(gNow);
            } else {
                dg = dg / (ng > 0 ? ng : 1.0);
            }
            change = dg;
            if (done > n && settled) {
                conv = true;
                break;
            }
        }
        prevMub = as.mub;
    }
    IFL_HIP(hipEventRecord(h->ev1, h->stream));
    IFL_HIP(hipEventSynchronize(h->ev1));
    float ms = 0;
    IFL_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    if (dJ_dmu) *dJ_dmu = as.mub;
    if (dJ_dprofile) {
        IFL_HIP(hipMemcpyAsync(dJ_dprofile, B.gprof[face], n_profile * 8, hipMemcpyDeviceToHost, h->stream));
        IFL_HIP(hipStreamSynchronize(h->stream));
    }
    if (report) {
        // objective at the current state
        double J = 0.0;
        std::vector<float> f(N);
        for (int q = 0; q < 4; ++q) {
            const bool has = q < 3 ? B.gu[q] != nullptr : B.gp != nullptr;
            if (!has) continue;
            const int fid = q < 3 ? IFL_FIELD_U + q : IFL_FIELD_P;
            if ((st = ifl_get_field(h, fid, f.data(), N)) != IFL_OK) return st;
            std::vector<float> wq(N);
            IFL_HIP(hipMemcpy(wq.data(), A.gw[q], N * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; ++i) J += (double)wq[i] * (double)f[i];
        }
        report->iterations = (int)done;
        report->converged = conv ? 1 : 0;
        report->J = J;
        report->final_scaled = as.scaled;
        report->elapsed_ms = ms;
        report->grad_change = change;
    }
    return IFL_OK;
}

extern "C" int ifl_adjoint_lambda(ifl_solver *h, int field, float *dst, size_t n) {
    if (!h || !dst || !h->adj) return fail(IFL_ERR_ARG, "no adjoint on this handle");
    const ifl::AdjBufs &B = h->adj->B;
    const float *src = nullptr;
    switch (field) {
        case IFL_FIELD_U: src = B.lu[0]; break;
        case IFL_FIELD_V: src = B.lu[1]; break;
        case IFL_FIELD_W: src = B.lu[2]; break;
        case IFL_FIELD_P: src = B.lp; break;
        case IFL_FIELD_UF: src = B.lf[0]; break;
        case IFL_FIELD_VF: src = B.lf[1]; break;
        case IFL_FIELD_WF: src = B.lf[2]; break;
        default: return fail(IFL_ERR_ARG, "bad adjoint field");
    }
    if (n != fieldSize(h->cfg, field)) return fail(IFL_ERR_ARG, "size mismatch");
    IFL_HIP(hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToHost, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    return IFL_OK;
}
