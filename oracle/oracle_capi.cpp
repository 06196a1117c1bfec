// oracle_capi.cpp - extern "C" surface of the CPU oracle (TEST INFRASTRUCTURE ONLY).
// Loaded with ctypes by the tests, the smoke check and the CPU-baseline leg of the benchmark.
// Values cross as double (exact for FP32 state), so float results compare bit for bit.
#include <omp.h>

#include <cstring>
#include <new>

#include "adjoint_oracle.hpp"
#include "simple_oracle.hpp"

namespace {

struct Handle {
    int fp64;
    oracle::Simple<float> *f;
    oracle::Simple<double> *d;
    oracle::Adjoint<float> *af = nullptr;
    oracle::Adjoint<double> *ad = nullptr;
};

// Flat layouts of adjoint vectors for ctypes: state = [u, v, w, p (N each), uf, vf, wf],
// parameters = [profiles of faces E, W, N, S, T, B (cells on each face), mu].
template <class Real>
size_t vecSize(const oracle::Adjoint<Real> &a) {
    return 4 * (size_t)a.Nc + a.s.uf.size() + a.s.vf.size() + a.s.wf.size();
}
template <class Real>
size_t parSize(const oracle::Adjoint<Real> &a) {
    size_t n = 1;
    for (int f = 0; f < 6; ++f) n += a.faceCells(f);
    return n;
}
template <class Real>
void unpackVec(const oracle::Adjoint<Real> &a, const double *src, oracle::AdjVec<Real> &x) {
    a.alloc(x);
    size_t o = 0;
    for (int d = 0; d < 3; ++d)
        for (auto &e : x.u[d]) e = (Real)src[o++];
    for (auto &e : x.p) e = (Real)src[o++];
    for (int d = 0; d < 3; ++d)
        for (auto &e : x.f[d]) e = (Real)src[o++];
}
template <class Real>
void packVec(const oracle::AdjVec<Real> &x, double *dst) {
    size_t o = 0;
    for (int d = 0; d < 3; ++d)
        for (auto e : x.u[d]) dst[o++] = (double)e;
    for (auto e : x.p) dst[o++] = (double)e;
    for (int d = 0; d < 3; ++d)
        for (auto e : x.f[d]) dst[o++] = (double)e;
}
template <class Real>
void unpackPar(const oracle::Adjoint<Real> &a, const double *src, oracle::AdjParams &t) {
    a.allocParams(t);
    size_t o = 0;
    for (int f = 0; f < 6; ++f)
        for (auto &e : t.prof[f]) e = src[o++];
    t.mu = src[o];
}
inline void packPar(const oracle::AdjParams &t, double *dst) {
    size_t o = 0;
    for (int f = 0; f < 6; ++f)
        for (auto e : t.prof[f]) dst[o++] = e;
    dst[o] = t.mu;
}

thread_local std::string g_err;

template <class Real>
int get_field(oracle::Simple<Real> *s, int id, double *dst, size_t n) {
    auto *v = s->field(id);
    if (!v) { g_err = "bad field id"; return IFL_ERR_ARG; }
    if (n != v->size()) { g_err = "size mismatch"; return IFL_ERR_ARG; }
    for (size_t q = 0; q < n; ++q) dst[q] = (double)(*v)[q];
    return IFL_OK;
}

template <class Real>
int set_field(oracle::Simple<Real> *s, int id, const double *src, size_t n) {
    auto *v = s->field(id);
    if (!v) { g_err = "bad field id"; return IFL_ERR_ARG; }
    if (n != v->size()) { g_err = "size mismatch"; return IFL_ERR_ARG; }
    for (size_t q = 0; q < n; ++q) (*v)[q] = (Real)src[q];
    // the oracle keeps separate u/v/w momentum matrices (Solver.h:28-30); setting MCOEF sets all
    if (id == IFL_FIELD_MCOEF) { s->vcoef = s->ucoef; s->wcoef = s->ucoef; }
    return IFL_OK;
}

template <class Real>
int run_stage(oracle::Simple<Real> *s, int stage, double *out, size_t n_out) {
    double tmp[8] = {0};
    switch (stage) {
        case IFL_STAGE_INIT_FACES: s->initFaceVel(); break;
        case IFL_STAGE_FACE_BC: s->applyBCsToFaceVel(); break;
        case IFL_STAGE_MOM_ASSEMBLE:   // Solver.cu:165-175: face BCs, coefficients, sources, BC rows
            s->applyBCsToFaceVel();
            s->assembleMomentum();
            break;
        case IFL_STAGE_MOM_RESIDUAL: s->momentumResidual(tmp); break;
        case IFL_STAGE_SMOOTH_U:
        case IFL_STAGE_SMOOTH_V:
        case IFL_STAGE_SMOOTH_W: {
            int q = stage - IFL_STAGE_SMOOTH_U;
            Real *fld = q == 0 ? s->u.data() : q == 1 ? s->v.data() : s->w.data();
            const Real *cf = q == 0 ? s->ucoef.data() : q == 1 ? s->vcoef.data() : s->wcoef.data();
            const Real *src = q == 0 ? s->bu.data() : q == 1 ? s->bv.data() : s->bw.data();
            tmp[0] = s->iterateMomentum(fld, cf, src, s->nIter[q], s->relax[q], s->tol[q]);
            tmp[1] = s->lastSweepSum;
            break;
        }
        case IFL_STAGE_SMOOTH_T:
            if (!s->energy) { g_err = "SMOOTH_T needs energy = 1"; return IFL_ERR_ARG; }
            tmp[0] = s->smoothTemperature();
            tmp[1] = s->lastSweepSum;
            break;
        case IFL_STAGE_RHIE_CHOW: s->RhieChowInterpolate(); break;
        case IFL_STAGE_PCORR_ASSEMBLE:
            s->calcPresCorrLinkCoef();
            tmp[0] = s->calcPresCorrSrcTerm();
            std::fill(s->pc.begin(), s->pc.end(), (Real)0);
            break;
        case IFL_STAGE_SMOOTH_P:
            tmp[0] = s->solvePressureCorrection();
            tmp[1] = s->lastSweepSum;
            break;
        case IFL_STAGE_UPDATE: s->updateField(tmp); break;
        case IFL_STAGE_OUTER: {
            ifl_iter_record rec;
            tmp[0] = s->outerIteration(rec) ? 1 : 0;
            tmp[1] = rec.scaled;
            break;
        }
        default: g_err = "bad stage"; return IFL_ERR_ARG;
    }
    if (out)
        for (size_t q = 0; q < n_out && q < 8; ++q) out[q] = tmp[q];
    return IFL_OK;
}

}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

// Planted mistakes (oracle::FAULT_*) for the tests that show a pin fails when the oracle is wrong.
void orc_set_fault(int code) { oracle::g_fault = code; }
int orc_get_fault(void) { return oracle::g_fault; }

void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}
int orc_get_threads(void) { return omp_get_max_threads(); }

int orc_create(const ifl_config *cfg, int fp64, void **out) {
    if (!cfg || !out) { g_err = "null argument"; return IFL_ERR_ARG; }
    Handle *h = new (std::nothrow) Handle{fp64, nullptr, nullptr};
    if (!h) return IFL_ERR_NOMEM;
    bool ok;
    if (fp64) { h->d = new oracle::Simple<double>(); ok = h->d->init(*cfg); if (!ok) g_err = h->d->error; }
    else { h->f = new oracle::Simple<float>(); ok = h->f->init(*cfg); if (!ok) g_err = h->f->error; }
    if (!ok) { delete h->f; delete h->d; delete h; return IFL_ERR_ARG; }
    *out = h;
    return IFL_OK;
}

int orc_destroy(void *hv) {
    Handle *h = (Handle *)hv;
    if (!h) return IFL_OK;
    delete h->af;
    delete h->ad;
    delete h->f;
    delete h->d;
    delete h;
    return IFL_OK;
}

size_t orc_field_size(void *hv, int id) {
    Handle *h = (Handle *)hv;
    if (h->fp64) { auto *v = h->d->field(id); return v ? v->size() : 0; }
    auto *v = h->f->field(id);
    return v ? v->size() : 0;
}

int orc_get_field(void *hv, int id, double *dst, size_t n) {
    Handle *h = (Handle *)hv;
    return h->fp64 ? get_field(h->d, id, dst, n) : get_field(h->f, id, dst, n);
}

int orc_set_field(void *hv, int id, const double *src, size_t n) {
    Handle *h = (Handle *)hv;
    return h->fp64 ? set_field(h->d, id, src, n) : set_field(h->f, id, src, n);
}

int orc_set_face_profile(void *hv, int face, const double *vals, size_t n) {
    Handle *h = (Handle *)hv;
    if (face < 0 || face > 5) { g_err = "bad face"; return IFL_ERR_ARG; }
    auto setp = [&](auto *s) {
        using R = typename std::remove_reference<decltype(s->u[0])>::type;
        size_t need = (face < 2) ? (size_t)s->ny * s->nz : (face < 4) ? (size_t)s->nx * s->nz : (size_t)s->nx * s->ny;
        if (!vals) { s->prof[face].clear(); return IFL_OK; }
        if (n != need) { g_err = "profile size mismatch"; return IFL_ERR_ARG; }
        s->prof[face].assign(n, (R)0);
        for (size_t q = 0; q < n; ++q) s->prof[face][q] = (R)vals[q];
        return IFL_OK;
    };
    return h->fp64 ? setp(h->d) : setp(h->f);
}

int orc_run_stage(void *hv, int stage, double *out, size_t n_out) {
    Handle *h = (Handle *)hv;
    return h->fp64 ? run_stage(h->d, stage, out, n_out) : run_stage(h->f, stage, out, n_out);
}

int orc_solve(void *hv, int max_iter, ifl_iter_record *hist, size_t cap, int *iters, int *converged) {
    Handle *h = (Handle *)hv;
    int st = IFL_OK;
    int n = h->fp64 ? h->d->solve(max_iter, hist, cap, converged, &st) : h->f->solve(max_iter, hist, cap, converged, &st);
    if (iters) *iters = n;
    return st;
}

// Overwrite the oracle's iteration counter / residual normalisation (warm starts).
int orc_set_iter(void *hv, long long it) {
    Handle *h = (Handle *)hv;
    if (h->fp64) h->d->iter = it; else h->f->iter = it;
    return IFL_OK;
}

// ---- discrete adjoint (adjoint_oracle.hpp) ----
// Set up the adjoint at the current state (which should be a converged fixed point).
// objective 0: pressure drop between the first and last cell layers along x; 1: linear
// weights given by orc_adj_set_weights.
int orc_adj_setup(void *hv, int objective) {
    Handle *h = (Handle *)hv;
    auto go = [&](auto *s, auto *&a) {
        using A = std::remove_pointer_t<std::remove_reference_t<decltype(a)>>;
        delete a;
        a = new A(*s);
        a->setup();
        if (objective == 0) a->setPressureDropObjective();
        return IFL_OK;
    };
    return h->fp64 ? go(h->d, h->ad) : go(h->f, h->af);
}

int orc_adj_sizes(void *hv, size_t *nvec, size_t *npar) {
    Handle *h = (Handle *)hv;
    if (h->fp64 ? !h->ad : !h->af) { g_err = "adjoint not set up"; return IFL_ERR_ARG; }
    if (h->fp64) { *nvec = vecSize(*h->ad); *npar = parSize(*h->ad); }
    else { *nvec = vecSize(*h->af); *npar = parSize(*h->af); }
    return IFL_OK;
}

// Objective weights as a state vector (u, v, w, p parts are used).
int orc_adj_set_weights(void *hv, const double *w) {
    Handle *h = (Handle *)hv;
    auto go = [&](auto *a) {
        using R = typename std::remove_reference_t<decltype(a->gp)>::value_type;
        oracle::AdjVec<R> x;
        unpackVec(*a, w, x);
        for (int d = 0; d < 3; ++d) a->gu[d] = x.u[d];
        a->gp = x.p;
        return IFL_OK;
    };
    return h->fp64 ? go(h->ad) : go(h->af);
}

double orc_adj_objective(void *hv) {
    Handle *h = (Handle *)hv;
    return h->fp64 ? h->ad->objective() : h->af->objective();
}

// out = G_x dx + G_theta dth (tangent of one outer iteration at the fixed point)
int orc_adj_tangent(void *hv, const double *dx, const double *dth, double *out) {
    Handle *h = (Handle *)hv;
    auto go = [&](auto *a) {
        using R = typename std::remove_reference_t<decltype(a->gp)>::value_type;
        oracle::AdjVec<R> x, o;
        oracle::AdjParams t;
        unpackVec(*a, dx, x);
        unpackPar(*a, dth, t);
        a->tangent(x, t, o);
        packVec(o, out);
        return IFL_OK;
    };
    return h->fp64 ? go(h->ad) : go(h->af);
}

// (out, tbar) = (G_x^T l, G_theta^T l), plus the objective gradient added to out
int orc_adj_transpose(void *hv, const double *l, double *out, double *tbar) {
    Handle *h = (Handle *)hv;
    auto go = [&](auto *a) {
        using R = typename std::remove_reference_t<decltype(a->gp)>::value_type;
        oracle::AdjVec<R> x, o;
        oracle::AdjParams t;
        unpackVec(*a, l, x);
        a->transpose(x, o, t);
        packVec(o, out);
        packPar(t, tbar);
        return IFL_OK;
    };
    return h->fp64 ? go(h->ad) : go(h->af);
}

// n adjoint iterations lambda <- G_x^T lambda + J_x^T; *scaled = the last scaled increment
int orc_adj_iterate(void *hv, int n, double *scaled) {
    Handle *h = (Handle *)hv;
    double sc = 0;
    for (int q = 0; q < n; ++q) sc = h->fp64 ? h->ad->iterate() : h->af->iterate();
    if (scaled) *scaled = sc;
    return IFL_OK;
}

int orc_adj_get_lambda(void *hv, double *dst) {
    Handle *h = (Handle *)hv;
    if (h->fp64) packVec(h->ad->lam, dst); else packVec(h->af->lam, dst);
    return IFL_OK;
}

int orc_adj_set_lambda(void *hv, const double *src) {
    Handle *h = (Handle *)hv;
    if (h->fp64) unpackVec(*h->ad, src, h->ad->lam); else unpackVec(*h->af, src, h->af->lam);
    return IFL_OK;
}

// G_theta^T lambda of the last iteration: the gradient once lambda has converged
int orc_adj_gradient(void *hv, double *tbar) {
    Handle *h = (Handle *)hv;
    packPar(h->fp64 ? h->ad->grad : h->af->grad, tbar);
    return IFL_OK;
}

// ---- p' multigrid internals (IFL_PSOLVE_MG), for the tests that pin it ----
// Build the levels from the current p' system (the p' matrix, b' and the momentum diagonal)
// with x = 0; *nlevels and dims[3 * l + {0,1,2}] receive the level count and sizes.
int orc_mg_build(void *hv, int *nlevels, int *dims) {
    Handle *h = (Handle *)hv;
    auto go = [&](auto *s) {
        s->mgBuild();
        *nlevels = (int)s->mg.size();
        for (size_t l = 0; l < s->mg.size(); ++l) {
            dims[3 * l] = s->mg[l].nx;
            dims[3 * l + 1] = s->mg[l].ny;
            dims[3 * l + 2] = s->mg[l].nz;
        }
        return IFL_OK;
    };
    return h->fp64 ? go(h->d) : go(h->f);
}
// which: 0 x, 1 b, 2 C, 3 E, 4 N, 5 T, 6 outlet term
int orc_mg_get(void *hv, int level, int which, double *dst, size_t n) {
    Handle *h = (Handle *)hv;
    auto go = [&](auto *s) {
        if (level < 0 || level >= (int)s->mg.size()) { g_err = "bad level"; return IFL_ERR_ARG; }
        auto &L = s->mg[level];
        auto *v = which == 0 ? &L.x : which == 1 ? &L.b : which == 2 ? &L.C : which == 3 ? &L.E : which == 4 ? &L.N
                 : which == 5 ? &L.T : &L.dg;
        if (n != v->size()) { g_err = "size mismatch"; return IFL_ERR_ARG; }
        for (size_t q = 0; q < n; ++q) dst[q] = (double)(*v)[q];
        return IFL_OK;
    };
    return h->fp64 ? go(h->d) : go(h->f);
}
int orc_mg_set(void *hv, int level, int which, const double *src, size_t n) {
    Handle *h = (Handle *)hv;
    auto go = [&](auto *s) {
        using R = typename std::remove_reference<decltype(s->u[0])>::type;
        if (level < 0 || level >= (int)s->mg.size()) { g_err = "bad level"; return IFL_ERR_ARG; }
        auto &L = s->mg[level];
        auto *v = which == 0 ? &L.x : which == 1 ? &L.b : nullptr;
        if (!v || n != v->size()) { g_err = "only x and b of a level can be set; size must match"; return IFL_ERR_ARG; }
        for (size_t q = 0; q < n; ++q) (*v)[q] = (R)src[q];
        return IFL_OK;
    };
    return h->fp64 ? go(h->d) : go(h->f);
}
// n V-cycles from the current level-0 x; *res receives the FP64 norm of b - A x on level 0.
int orc_mg_cycles(void *hv, int n, double *res) {
    Handle *h = (Handle *)hv;
    auto go = [&](auto *s) {
        if (s->mg.empty()) { g_err = "no levels: call orc_mg_build first"; return IFL_ERR_ARG; }
        for (int q = 0; q < n; ++q) s->mgCycle(0);
        auto &L = s->mg[0];
        double t = 0;
        for (int k = 0; k < L.nz; ++k)
            for (int j = 0; j < L.ny; ++j)
                for (int i = 0; i < L.nx; ++i) {
                    const double r = s->mgResidual(L, i, j, k);
                    t += r * r;
                }
        if (res) *res = std::sqrt(t);
        return IFL_OK;
    };
    return h->fp64 ? go(h->d) : go(h->f);
}
// n smoothing sweeps of one level (a smoother pin)
int orc_mg_smooth(void *hv, int level, int n) {
    Handle *h = (Handle *)hv;
    auto go = [&](auto *s) {
        if (level < 0 || level >= (int)s->mg.size()) { g_err = "bad level"; return IFL_ERR_ARG; }
        s->mgSmooth(s->mg[level], n);
        return IFL_OK;
    };
    return h->fp64 ? go(h->d) : go(h->f);
}

size_t orc_config_size(void) { return sizeof(ifl_config); }
size_t orc_record_size(void) { return sizeof(ifl_iter_record); }

}  // extern "C"
