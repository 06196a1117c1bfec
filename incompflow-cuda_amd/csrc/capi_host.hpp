// capi_host.hpp - the host runtime shared by the C-ABI sources (capi*.hip): the handle, its
// device arena, the launch plan of one outer iteration, graph capture and the monitor reads.
// Internal to libincompflow.so; nothing here crosses the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/incompflow.h"
#include "ifl_adjoint.hpp"
#include "ifl_device.hpp"
#include "ifl_kernels.hpp"
#include "ifl_mg.hpp"

namespace ifl_host {

// Thread-local description of the last failure (ifl_last_error).
extern thread_local std::string g_err;

inline int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define IFL_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::ifl_host::fail(IFL_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kHistCap = 4096;
constexpr int kMaxSweeps = 100000;

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace ifl_host

// Device levels of the multigrid p' solve of one handle (IFL_PSOLVE_MG), made when first needed.
struct MgHost {
    char *arena = nullptr;        // levels 1.. (x, b, C, E, N(, T), outlet term) and the 2D level table
    ifl::MgLaunch launch{};       // 2D: launch.dev points into the arena
    ifl::MgLaunch3 launch3{};     // 3D
};

struct ifl_solver {
    ifl_config cfg{};
    ifl::CaseParams P{};
    hipStream_t stream = nullptr;
    char *arena = nullptr;
    size_t arenaBytes = 0;
    float *prof[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t graph = nullptr;
    bool graphDirty = true;
    long long iter = 0;   // host mirror of CaseState::iter
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float tolf[5];
    int tile = 0;     // 2D sweep tile variant
    struct AdjHost *adj = nullptr;   // discrete adjoint buffers, made by the first ifl_adjoint call
    MgHost *mg = nullptr;            // multigrid levels of the p' solve, made when first needed
};

// Device buffers of the adjoint of one handle and the graph of one adjoint iteration.
struct AdjHost {
    char *arena = nullptr;
    ifl::AdjBufs B{};
    float *gw[4] = {nullptr, nullptr, nullptr, nullptr};   // objective weights u, v, w, p
    hipGraphExec_t graph = nullptr;
};

namespace ifl_host {

int validate(const ifl_config &c);
void fillParams(ifl_solver *h);
int partStride(const ifl::CaseParams &P);

// What one outer iteration launches, for one case or a batch of cases with the same settings.
struct Plan {
    ifl::Lc L;
    const ifl_config *cfg;   // sweep counts, smoother, refinement, sweep block (shared by a batch)
    int dim;
    const ifl::MgLaunch *mg = nullptr;   // multigrid levels of the p' solve (one case, 2D)
    const ifl::MgLaunch3 *mg3 = nullptr; // the same, 3D
};

int sweepBlockOf(const ifl_config &c, int dim, int slot);
bool momentumStartsInPartner(const ifl_config &c, int dim);
bool pressureStartsInPartner(const ifl_config &c, int dim);
int slotEq(int slot);
int slotSweeps(const ifl_config &c, int slot);
int sweepKind(const Plan &pl, int slot);
int sweepRecordsSingle(const Plan &pl, int slot);

// How a smoother call is closed.  The outer-iteration path (fast) lets a call with an odd launch
// count start in the partner buffer (which the kernels keep equal to the field: the update writes
// the velocities to both, p' and its correction start from zero in both), so the result lands in
// the caller's field without a copy, and hands the call's closing step to workgroup 0 of the next
// kernel (pending).  The stage-by-stage path copies back and launches k_smooth_finish, and leaves
// the momentum fields' partners equal to them.
struct SmoothMode {
    bool fast = false;
    ifl::Finish *pending = nullptr;
    bool accumLast = false;      // the last launch adds the correction into p' (2D red-black)
};

void enqueueSmoothers(const Plan &pl, const int *slots, int neq, const SmoothMode &md = SmoothMode{});
void enqueueMomentumSmoothers(const Plan &pl, const SmoothMode &md = SmoothMode{});
void enqueueTemperatureSmoother(const Plan &pl, const SmoothMode &md = SmoothMode{});
ifl::Finish enqueuePressureCorrection(const Plan &pl, bool fast = false);
void enqueueOuter(const Plan &pl);
Plan planOf(ifl_solver *h);
int captureGraph(hipStream_t stream, const Plan &pl, hipGraphExec_t *out);
int buildGraph(ifl_solver *h);
// Allocate the multigrid levels when the configuration asks for IFL_PSOLVE_MG, and copy its
// settings into the launch description (before any capture or launch of a p' solve).
int mgEnsure(ifl_solver *h);
int readState(ifl_solver *h, ifl::CaseState &s);
void printLog(const ifl_iter_record &r, long long it, int dim);
int setDone(ifl_solver *h, int v);
int armCase(ifl_solver *h, hipStream_t s, long long cap);
int copyHistory(ifl_solver *h, hipStream_t s, long long a, long long b, ifl_iter_record *dst);

// Host bookkeeping of a solve over the records of iterations [a, b) (it0: first iteration of
// the solve): history output, sweep count, log lines (Solver.cu:270-285).
struct SolveTally {
    long long it0 = 0;
    int maxIter = 0;
    double sweeps = 0.0;
    ifl_iter_record *history = nullptr;
    size_t histCap = 0;
};
void tally(ifl_solver *h, SolveTally &t, const ifl_iter_record *recs, long long a, long long b, bool doneAtEnd,
           int logEvery);
// The host loop of a solve: outer iterations in chunks of check_every (64 when 0), each chunk
// enqueued by launch(n) on h->stream; after each chunk the monitor is read and the new records go,
// ringCap at a time through ring, to tally.  Stops at max_iter or when the device sets its stop flag.
template <class Launch>
int runChunks(ifl_solver *h, int max_iter, Launch &&launch, SolveTally &t, ifl_iter_record *ring, int ringCap,
              ifl::CaseState &s) {
    const int chunk = h->cfg.check_every > 0 ? std::min(h->cfg.check_every, kHistCap) : 64;
    long long launched = 0;
    int st;
    while (launched < max_iter) {
        const int n = (int)std::min<long long>(chunk, max_iter - launched);
        if ((st = launch(n)) != IFL_OK) return st;
        launched += n;
        if ((st = readState(h, s)) != IFL_OK) return st;
        for (long long a = h->iter; a < s.iter; a += ringCap) {
            const long long b = std::min<long long>(s.iter, a + ringCap);
            if ((st = copyHistory(h, h->stream, a, b, ring)) != IFL_OK) return st;
            IFL_HIP(hipStreamSynchronize(h->stream));
            tally(h, t, ring, a, b, s.done != 0 && b == s.iter, h->cfg.log_every);
        }
        h->iter = s.iter;
        if (s.done) break;
    }
    return IFL_OK;
}

// The report and return status of a solve that ran from iteration it0 (ifl_solve's contract).
int finishSolve(ifl_solver *h, long long it0, const SolveTally &t, const ifl::CaseState &s, double ms,
                ifl_report *report);
float *fieldPtr(ifl_solver *h, int id);
// Lay out the arena of a case (every device buffer but the face profiles and the multigrid levels)
// from base, or only size it when base is null; returns its bytes.  external[q] (u, v, w, p, temp;
// external may be null) marks a cell field that the caller owns: it takes no space and its pointer
// is left to the caller.  Sets P.partStride; fillParams must have run.
size_t arenaLayout(ifl_solver *h, char *base, const bool *external);
// Lay out the multigrid levels 1.. of a 2D case and their device table from base (null: size only).
size_t mgLayout(const ifl_config &c, MgHost &m, char *base);
// Copy the multigrid settings of a configuration into its launch description.
void mgSettings(const ifl_config &c, MgHost &m);
size_t fieldSize(const ifl_config &c, int id);

// A copy of the whole device arena (fields, matrices, monitor, history) and the iteration count.
struct Snapshot {
    char *buf = nullptr;
    size_t bytes = 0;
    long long iter = 0;
    ~Snapshot() {
        if (buf) (void)hipFree(buf);
    }
};

int snapTake(ifl_solver *h, Snapshot &s);
int snapRestore(ifl_solver *h, const Snapshot &s);

}  // namespace ifl_host
