// capi_noalloc.hip - ifl_solve_noalloc: a solve that allocates no device memory, for callers that
// cannot call hipMalloc (serving loops, harnesses that forbid allocation).
//
// The case's arena (every buffer of ifl_create's arena except the caller's cell fields) and its
// multigrid levels are carved from `scratch`, a file-scope device array of IFL_NOALLOC_SCRATCH_BYTES
// that every GPU the library runs on holds until the process ends.  The kernels are the ones
// ifl_solve replays from its graph, launched directly on the caller's stream, so the fields equal
// those of ifl_create + ifl_set_field + ifl_solve bit for bit.  Calls on one GPU are serialised by
// a per-device lock held for the whole call (each call ends with its stream drained); every call
// clears the part of the scratch it uses before its first kernel.
#include <chrono>
#include <mutex>

#include "capi_host.hpp"

namespace {

static __device__ __align__(256) char scratch[IFL_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_devLock[kMaxDevices];
char *g_scratch[kMaxDevices];      // device address of `scratch` on each device (null: not looked up)

__global__ void k_fill_words(uint32_t *dst, size_t n, uint32_t v) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = v;
}

void fillWords(void *dst, size_t words, uint32_t v, hipStream_t s) {
    if (!words) return;
    const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_words, dim3(blocks), dim3(256), 0, s, (uint32_t *)dst, words, v);
}

uint32_t floatBits(double v) {
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

}  // namespace

using namespace ifl_host;

extern "C" int ifl_solve_noalloc(const ifl_config *cfg, void *stream, float *u, float *v, float *w, float *p,
                                 float *temp, const float *const *face_profiles, int max_iter,
                                 ifl_report *report, ifl_iter_record *history, size_t hist_cap) {
    if (report) std::memset(report, 0, sizeof(*report));
    if (!cfg || !u || !v || !p) return fail(IFL_ERR_ARG, "null argument (cfg, u, v and p are required)");
    int st = validate(*cfg);
    if (st != IFL_OK) return st;
    if (cfg->dim == 3 && !w) return fail(IFL_ERR_ARG, "w is required in 3D");
    if (cfg->energy && !temp) return fail(IFL_ERR_ARG, "temp is required with energy = 1");
    if (max_iter <= 0) max_iter = cfg->num_outer_iter;

    // Host description of the case (no device memory of its own).
    ifl_solver h{};
    h.cfg = *cfg;
    if (face_profiles)
        for (int f = 0; f < 6; ++f) h.prof[f] = const_cast<float *>(face_profiles[f]);
    fillParams(&h);
    h.tile = (cfg->dim == 2) ? (cfg->tile ? cfg->tile : ifl::default_tile(cfg->nx, cfg->ny)) : 0;
    const bool external[5] = {true, true, w != nullptr, true, temp != nullptr};
    const size_t arenaBytes = arenaLayout(&h, nullptr, external);
    MgHost mg{};
    const bool useMg = cfg->p_solver == IFL_PSOLVE_MG;
    const size_t mgBytes = useMg ? mgLayout(*cfg, mg, nullptr) : 0;
    const size_t need = arenaBytes + mgBytes;
    if (need > (size_t)IFL_NOALLOC_SCRATCH_BYTES)
        return fail(IFL_ERR_ARG, "ifl_solve_noalloc: the case needs " + std::to_string(need)
                                     + " bytes of device scratch; the library holds "
                                     + std::to_string((size_t)IFL_NOALLOC_SCRATCH_BYTES) + " per GPU");

    hipStream_t s = (hipStream_t)stream;
    int dev = 0;
    IFL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(IFL_ERR_DEVICE, "device ordinal beyond the scratch table");
    if (s) {
        hipDevice_t sdev = 0;
        IFL_HIP(hipStreamGetDevice(s, &sdev));
        if ((int)sdev != dev) return fail(IFL_ERR_ARG, "the stream belongs to another device than the current one");
    }
    std::lock_guard<std::mutex> lock(g_devLock[dev]);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        IFL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (char *)ptr;
    }
    char *base = g_scratch[dev];

    // Carve the arena and the levels, then start the case as ifl_create does, from the caller's fields.
    h.stream = s;
    h.arena = base;
    h.arenaBytes = arenaBytes;
    arenaLayout(&h, base, external);
    ifl::CaseParams &P = h.P;
    P.u = u;
    P.v = v;
    P.p = p;
    if (w) P.w = w;
    if (temp) P.temp = temp;
    if (useMg) {
        mgLayout(*cfg, mg, base + arenaBytes);
        mgSettings(*cfg, mg);
        h.mg = &mg;
    }
    const size_t N = (size_t)P.N;
    fillWords(base, need / 4, 0u, s);
    if (!w) fillWords(P.w, N, floatBits(cfg->init[2]), s);
    if (!temp) fillWords(P.temp, N, floatBits(cfg->init[4]), s);
    IFL_HIP(hipGetLastError());
    // u, v, w also into their smoothers' partner buffers, which stay equal to them
    IFL_HIP(hipMemcpyAsync(P.alt[0], P.u, N * 4, hipMemcpyDeviceToDevice, s));
    IFL_HIP(hipMemcpyAsync(P.alt[1], P.v, N * 4, hipMemcpyDeviceToDevice, s));
    IFL_HIP(hipMemcpyAsync(P.alt[2], P.w, N * 4, hipMemcpyDeviceToDevice, s));
    if (useMg && cfg->dim == 2)
        IFL_HIP(hipMemcpyAsync(base + arenaBytes, mg.launch.lv, sizeof(ifl::MgLevel) * ifl::kMgMaxLevels,
                               hipMemcpyHostToDevice, s));
    IFL_HIP(hipStreamSynchronize(s));

    if ((st = armCase(&h, s, 0)) != IFL_OK) return st;
    const auto t0 = std::chrono::steady_clock::now();
    SolveTally t;
    t.it0 = 0;
    t.maxIter = max_iter;
    t.history = history;
    t.histCap = hist_cap;
    const Plan pl = planOf(&h);
    ifl::launch_init_faces(pl.L);                                    // Solver.cu:162-164
    auto eager = [&pl](int n) -> int {
        for (int q = 0; q < n; ++q) enqueueOuter(pl);
        IFL_HIP(hipGetLastError());
        return IFL_OK;
    };
    constexpr int kRing = 64;
    ifl_iter_record ring[kRing];
    ifl::CaseState cs{};
    if ((st = runChunks(&h, max_iter, eager, t, ring, kRing, cs)) != IFL_OK) return st;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return finishSolve(&h, 0, t, cs, ms, report);
}
