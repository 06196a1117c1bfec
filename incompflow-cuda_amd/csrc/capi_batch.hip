// capi_batch.hip - batched solves over uneven cases (ifl_batch_*).
#include "capi_host.hpp"

using namespace ifl_host;

// ---------------- batched solves ----------------
// The batch owns one solver handle per case (arena, monitor, history ring) and runs them all
// through one graph on its own stream.  A device table holds every case's CaseParams and, per
// workgroup kind, the prefix sums of the cases' workgroup counts; a kernel launch covers the sum
// and each workgroup finds its case by binary search.  When half the cases of the graph have
// stopped, the table and graph are rebuilt over the cases still running.
struct ifl_batch {
    std::vector<ifl_solver *> cases;
    ifl_config cfg{};                 // the settings that fix the launch sequence (from case 0)
    int tile = 0;
    hipStream_t stream = nullptr;
    char *table = nullptr;            // device: CaseParams[n], then start/dimx per kind
    size_t tableBytes = 0;
    std::vector<char> hostTable;
    hipGraphExec_t graph = nullptr;
    std::vector<int> inGraph;         // case indices the graph covers
    ifl::Lc L{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

int batchValidate(const ifl_config *c, int n) {
    for (int i = 0; i < n; ++i) {
        int st = validate(c[i]);
        if (st != IFL_OK) return fail(st, "case " + std::to_string(i) + ": " + g_err);
        if (c[i].p_solver != IFL_PSOLVE_SWEEPS)
            return fail(IFL_ERR_ARG, "case " + std::to_string(i) + ": a batch runs the p' sweeps (p_solver = IFL_PSOLVE_SWEEPS)");
        const ifl_config &a = c[0], &b = c[i];
        bool same = a.dim == b.dim && a.smoother == b.smoother && a.p_refine == b.p_refine
                    && a.sweep_block == b.sweep_block && a.sweep_block_mom == b.sweep_block_mom
                    && a.tile == b.tile && a.energy == b.energy
                    && (!a.energy || a.n_iter_t == b.n_iter_t);
        for (int q = 0; q < 4; ++q) same = same && a.n_iter[q] == b.n_iter[q];
        if (!same)
            return fail(IFL_ERR_ARG, "case " + std::to_string(i)
                                         + ": dim, smoother, n_iter, p_refine, sweep_block(_mom), tile, energy and n_iter_t "
                                           "must agree across a batch");
    }
    return IFL_OK;
}

// Tile of a batch: the configured one, else the largest whose total tile count fills the GPU.
int batchTile(const ifl_batch *b) {
    if (b->cfg.dim != 2) return 0;
    if (b->cfg.tile) return b->cfg.tile;
    for (int t = ifl::TILE_64x16; t < ifl::TILE_32x4; ++t) {
        long long tot = 0;
        for (ifl_solver *h : b->cases) tot += ifl::tile_count(h->P.nx, h->P.ny, t);
        if (tot >= 768) return t;
    }
    return ifl::TILE_32x4;
}

// Upload the table for the cases `ids` and capture the outer-iteration graph over them.
int batchBuild(ifl_batch *b, const std::vector<int> &ids) {
    if (b->graph) {
        IFL_HIP(hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
    }
    const int n = (int)ids.size();
    const size_t pBytes = align256(sizeof(ifl::CaseParams) * n);
    const size_t kBytes = sizeof(int) * (2 * n + 1);
    // per-workgroup tables: count the workgroups of every kind first
    size_t blkBytes[ifl::MAP_KINDS];
    for (int k = 0; k < ifl::MAP_KINDS; ++k) {
        long long tot = 0;
        for (int c = 0; c < n; ++c) {
            const ifl::CaseParams &P = b->cases[ids[c]]->P;
            switch (k) {
                case ifl::MAP_CELL: tot += ifl::cell_blocks(P); break;
                case ifl::MAP_TILE: tot += P.dim == 2 ? ifl::tile_count(P.nx, P.ny, b->tile ? b->tile : ifl::TILE_64x16) : 0; break;
                case ifl::MAP_COPY: tot += ifl::copy_blocks(P); break;
                case ifl::MAP_ZMARCH: tot += P.dim == 3 ? ifl::zmarch_blocks(P, false) : 0; break;
                case ifl::MAP_ZMARCH_P: tot += P.dim == 3 ? ifl::zmarch_blocks(P, true) : 0; break;
                default: tot += 1; break;
            }
        }
        blkBytes[k] = align256(sizeof(int4) * (size_t)std::max(1LL, tot));
    }
    size_t need = pBytes + ifl::MAP_KINDS * align256(kBytes);
    for (int k = 0; k < ifl::MAP_KINDS; ++k) need += blkBytes[k];
    if (need > b->tableBytes) {
        if (b->table) IFL_HIP(hipFree(b->table));
        b->table = nullptr;
        b->tableBytes = 0;
        IFL_HIP(hipMalloc((void **)&b->table, need));
        b->tableBytes = need;
    }
    std::vector<char> &ht = b->hostTable;
    ht.assign(need, 0);
    ifl::CaseParams *hp = (ifl::CaseParams *)ht.data();
    ifl::Lc &L = b->L;
    L = ifl::Lc{};
    L.batch = true;
    L.tile = b->tile ? b->tile : ifl::TILE_64x16;
    L.s = b->stream;
    L.P = &b->cases[ids[0]]->P;
    L.M.ncase = n;
    L.M.P = (const ifl::CaseParams *)b->table;
    for (int k = 0; k < ifl::MAP_KINDS; ++k) {
        const size_t off = pBytes + k * align256(kBytes);
        int *start = (int *)(ht.data() + off), *dimx = start + n + 1;
        long long acc = 0;
        for (int c = 0; c < n; ++c) {
            const ifl::CaseParams &P = b->cases[ids[c]]->P;
            int cnt = 1, dx = 1;
            switch (k) {
                case ifl::MAP_CELL: cnt = ifl::cell_blocks(P); dx = ifl::cell_dimx(P); break;
                case ifl::MAP_TILE:
                    cnt = P.dim == 2 ? ifl::tile_count(P.nx, P.ny, L.tile) : 0;
                    dx = P.dim == 2 ? ifl::tile_dimx(P.nx, L.tile) : 1;
                    break;
                case ifl::MAP_COPY: cnt = ifl::copy_blocks(P); dx = cnt; break;
                case ifl::MAP_ZMARCH: cnt = P.dim == 3 ? ifl::zmarch_blocks(P, false) : 0; dx = ifl::zmarch_dimx(P, false); break;
                case ifl::MAP_ZMARCH_P: cnt = P.dim == 3 ? ifl::zmarch_blocks(P, true) : 0; dx = ifl::zmarch_dimx(P, true); break;
                default: break;
            }
            start[c] = (int)acc;
            dimx[c] = dx;
            acc += cnt;
        }
        if (acc >= (1LL << 31)) return fail(IFL_ERR_ARG, "batch too large for one launch");
        start[n] = (int)acc;
        L.total[k] = (int)acc;
        L.M.start[k] = (const int *)(b->table + off);
        L.M.dimx[k] = (const int *)(b->table + off) + n + 1;
    }
    size_t boff = pBytes + ifl::MAP_KINDS * align256(kBytes);
    for (int k = 0; k < ifl::MAP_KINDS; ++k) {
        const int *start = (const int *)(ht.data() + pBytes + k * align256(kBytes));
        const int *dimx = start + n + 1;
        int4 *e = (int4 *)(ht.data() + boff);
        for (int c = 0; c < n; ++c)
            for (int w = start[c]; w < start[c + 1]; ++w) e[w] = make_int4(c, w - start[c], start[c + 1] - start[c], dimx[c]);
        L.M.blk[k] = (const int4 *)(b->table + boff);
        boff += blkBytes[k];
    }
    for (int c = 0; c < n; ++c) hp[c] = b->cases[ids[c]]->P;
    IFL_HIP(hipMemcpyAsync(b->table, ht.data(), need, hipMemcpyHostToDevice, b->stream));
    IFL_HIP(hipStreamSynchronize(b->stream));
    Plan pl;
    pl.L = L;
    pl.cfg = &b->cfg;
    pl.dim = b->cfg.dim;
    const int st = captureGraph(b->stream, pl, &b->graph);
    if (st != IFL_OK) return st;
    b->inGraph = ids;
    return IFL_OK;
}

}  // namespace

extern "C" {

int ifl_batch_create(const ifl_config *cfgs, int n, ifl_batch **out) {
    if (!cfgs || !out || n <= 0) return fail(IFL_ERR_ARG, "null argument or empty batch");
    *out = nullptr;
    int st = batchValidate(cfgs, n);
    if (st != IFL_OK) return st;
    ifl_batch *b = new (std::nothrow) ifl_batch();
    if (!b) return fail(IFL_ERR_NOMEM, "out of host memory");
    b->cfg = cfgs[0];
    for (int i = 0; i < n; ++i) {
        ifl_solver *h = nullptr;
        st = ifl_create(&cfgs[i], &h);
        if (st != IFL_OK) {
            std::string m = "case " + std::to_string(i) + ": " + g_err;
            ifl_batch_destroy(b);
            return fail(st, m);
        }
        b->cases.push_back(h);
    }
    b->tile = batchTile(b);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        std::string m = std::string("batch stream: ") + hipGetErrorString(e);
        ifl_batch_destroy(b);
        return fail(IFL_ERR_DEVICE, m);
    }
    *out = b;
    return IFL_OK;
}

int ifl_batch_destroy(ifl_batch *b) {
    if (!b) return IFL_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->table) (void)hipFree(b->table);
    for (ifl_solver *h : b->cases) ifl_destroy(h);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return IFL_OK;
}

int ifl_batch_size(const ifl_batch *b) { return b ? (int)b->cases.size() : 0; }

int ifl_batch_case(ifl_batch *b, int i, ifl_solver **out) {
    if (!b || !out || i < 0 || i >= (int)b->cases.size()) return fail(IFL_ERR_ARG, "bad batch or case index");
    *out = b->cases[i];
    return IFL_OK;
}

namespace {
// Pinned host buffers of the batch readback: two chunk slots, each with every case's state and the
// iteration records of one chunk, and the event that marks their arrival.  Released after the
// stream has drained (an error return may leave copies into them in flight).
struct PinnedChunks {
    hipStream_t stream = nullptr;
    struct Slot {
        ifl::CaseState *state = nullptr;
        ifl_iter_record *rec = nullptr;
        hipEvent_t ev = nullptr;
        std::vector<int> ids;
        std::vector<long long> first;
    } slot[2];
    int init(hipStream_t s, int n, int chunk) {
        stream = s;
        for (Slot &S : slot) {
            IFL_HIP(hipHostMalloc((void **)&S.state, sizeof(ifl::CaseState) * (size_t)n, hipHostMallocDefault));
            IFL_HIP(hipHostMalloc((void **)&S.rec, sizeof(ifl_iter_record) * (size_t)n * (size_t)chunk,
                                  hipHostMallocDefault));
            IFL_HIP(hipEventCreateWithFlags(&S.ev, hipEventDisableTiming));
        }
        return IFL_OK;
    }
    ~PinnedChunks() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (Slot &S : slot) {
            if (S.ev) (void)hipEventDestroy(S.ev);
            if (S.rec) (void)hipHostFree(S.rec);
            if (S.state) (void)hipHostFree(S.state);
        }
    }
};
}  // namespace

int ifl_batch_solve(ifl_batch *b, int max_iter, ifl_report *reports) {
    if (!b) return fail(IFL_ERR_ARG, "null batch");
    const int n = (int)b->cases.size();
    int limit = max_iter;
    if (limit <= 0)
        for (ifl_solver *h : b->cases) limit = std::max(limit, h->cfg.num_outer_iter);
    if (reports) std::memset(reports, 0, sizeof(ifl_report) * n);
    int st;
    for (ifl_solver *h : b->cases) IFL_HIP(hipStreamSynchronize(h->stream));
    // every case runs at most its own limit (num_outer_iter, or max_iter when given)
    std::vector<long long> it0(n);
    std::vector<SolveTally> t(n);
    for (int i = 0; i < n; ++i) {
        ifl_solver *h = b->cases[i];
        const int own = max_iter > 0 ? max_iter : h->cfg.num_outer_iter;
        it0[i] = h->iter;
        t[i].it0 = h->iter;
        t[i].maxIter = own;
        if ((st = armCase(h, b->stream, h->iter + own)) != IFL_OK) return st;
    }
    std::vector<int> live(n);
    for (int i = 0; i < n; ++i) live[i] = i;
    if ((st = batchBuild(b, live)) != IFL_OK) return st;
    // iterations between host reads; the pinned readback holds two chunks of every case's records,
    // so the chunk shrinks to keep that under 256 MiB (it changes only how soon the host notices)
    const size_t perIter = 2 * sizeof(ifl_iter_record) * (size_t)n;
    const int chunkCap = (int)std::max<size_t>(1, ((size_t)256 << 20) / perIter);
    const int chunk = std::min(b->cfg.check_every > 0 ? std::min(b->cfg.check_every, kHistCap) : 64, chunkCap);
    IFL_HIP(hipEventRecord(b->ev0, b->stream));
    for (ifl_solver *h : b->cases)
        if (h->iter == 0) {                                           // Solver.cu:162-164
            Plan pl = planOf(h);
            pl.L.s = b->stream;
            ifl::launch_init_faces(pl.L);
        }
    std::vector<int> finalDone(n, 0);
    std::vector<double> finalScaled(n, 0.0);
    // Pipelined chunks: chunk j+1 is enqueued before chunk j's states and iteration records are
    // read, and both go to pinned host memory in stream order right behind chunk j, so the device
    // never waits for this bookkeeping.  A case that stops during chunk j runs no further iteration
    // on the device (its workgroups return at once), so the chunk launched ahead of that knowledge
    // changes nothing of its result.  A running case executes every iteration of a chunk, so the
    // first record of its next chunk is known before the readback.
    PinnedChunks pc;
    if ((st = pc.init(b->stream, n, chunk)) != IFL_OK) return st;
    std::vector<long long> next(n);          // first iteration of each case's next chunk
    for (int i = 0; i < n; ++i) next[i] = b->cases[i]->iter;
    std::vector<char> stopped(n, 0);
    long long launched = 0;
    auto enqueueChunk = [&](int slot) -> int {
        PinnedChunks::Slot &S = pc.slot[slot];
        const int k = (int)std::min<long long>(chunk, limit - launched);
        for (int q = 0; q < k; ++q) IFL_HIP(hipGraphLaunch(b->graph, b->stream));
        launched += k;
        S.ids = live;
        S.first.resize(live.size());
        for (size_t c = 0; c < live.size(); ++c) {
            const int i = live[c];
            ifl_solver *h = b->cases[i];
            IFL_HIP(hipMemcpyAsync(&S.state[c], h->P.st, sizeof(ifl::CaseState), hipMemcpyDeviceToHost, b->stream));
            int r = copyHistory(h, b->stream, next[i], next[i] + k, S.rec + c * (size_t)chunk);
            if (r != IFL_OK) return r;
            S.first[c] = next[i];
            next[i] += k;
        }
        IFL_HIP(hipEventRecord(S.ev, b->stream));
        return IFL_OK;
    };
    auto readChunk = [&](int slot) -> int {
        PinnedChunks::Slot &S = pc.slot[slot];
        IFL_HIP(hipEventSynchronize(S.ev));
        for (size_t c = 0; c < S.ids.size(); ++c) {
            const int i = S.ids[c];
            if (stopped[i]) continue;                 // stopped in an earlier chunk
            ifl_solver *h = b->cases[i];
            const ifl::CaseState &cs = S.state[c];
            const long long a = h->iter, e = cs.iter;
            if (a != S.first[c] || e < a || e - a > chunk)
                return fail(IFL_ERR_DEVICE, "batch: case " + std::to_string(i) + " ran an unexpected number of iterations");
            if (e > a) tally(h, t[i], S.rec + c * (size_t)chunk, a, e, cs.done != 0, 0);
            h->iter = e;
            finalDone[i] = cs.done;
            finalScaled[i] = cs.lastScaled;
            if (cs.done) stopped[i] = 1;
        }
        std::vector<int> still;
        for (int i : live)
            if (!stopped[i]) still.push_back(i);
        live.swap(still);
        return IFL_OK;
    };
    int cur = 0;
    bool pending[2] = {false, false};
    if (launched < limit && !live.empty()) {
        if ((st = enqueueChunk(cur)) != IFL_OK) return st;
        pending[cur] = true;
    }
    while (pending[cur]) {
        const int nx = cur ^ 1;
        if (launched < limit && !live.empty()) {
            if ((st = enqueueChunk(nx)) != IFL_OK) return st;
            pending[nx] = true;
        }
        if ((st = readChunk(cur)) != IFL_OK) return st;
        pending[cur] = false;
        if (!live.empty() && launched < limit && 2 * live.size() <= b->inGraph.size()) {
            // shrink the graph to the cases still running, after the chunk in flight has ended
            if (pending[nx]) {
                if ((st = readChunk(nx)) != IFL_OK) return st;
                pending[nx] = false;
            }
            if (!live.empty() && launched < limit) {
                if ((st = batchBuild(b, live)) != IFL_OK) return st;
                if ((st = enqueueChunk(nx)) != IFL_OK) return st;
                pending[nx] = true;
            }
        }
        cur = nx;
    }
    IFL_HIP(hipEventRecord(b->ev1, b->stream));
    IFL_HIP(hipEventSynchronize(b->ev1));
    float ms = 0;
    IFL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    bool diverged = false;
    for (int i = 0; i < n; ++i) {
        if ((st = armCase(b->cases[i], b->stream, 0)) != IFL_OK) return st;
        if (finalDone[i] == 3) finalDone[i] = 0;
        diverged = diverged || finalDone[i] == 2;
        if (reports) {
            ifl_report &r = reports[i];
            r.iterations = (int)(b->cases[i]->iter - it0[i]);
            r.converged = finalDone[i] == 1;
            r.status = finalDone[i] == 2 ? IFL_ERR_DIVERGED : IFL_OK;
            r.elapsed_ms = ms;
            r.final_scaled = finalScaled[i];
            r.sweeps_total = t[i].sweeps;
        }
    }
    if (diverged) return fail(IFL_ERR_DIVERGED, "non-finite residual: a case of the batch diverged");
    for (int i = 0; i < n; ++i)
        if (b->cases[i]->cfg.fail_if_not_converged && finalDone[i] != 1)
            return fail(IFL_ERR_NOT_CONVERGED, "case " + std::to_string(i) + " did not meet its stop rule");
    return IFL_OK;
}


// The batch's launch settings by measurement: every (tile, sweep_block, sweep_block_mom) of the
// candidates runs trial_iters outer iterations of the whole batch from its current state (after 3
// warm-up iterations), every case is restored afterwards, and the fastest per iteration is kept.
// The fixed point of each case does not depend on these settings; n_iter is left as configured.
int ifl_batch_autotune(ifl_batch *b, int trial_iters, ifl_tune_result *out) {
    if (!b || !out) return fail(IFL_ERR_ARG, "null argument");
    std::memset(out, 0, sizeof(*out));
    const int n = (int)b->cases.size();
    const int trial = trial_iters > 0 ? trial_iters : 30;
    out->tile = b->tile;
    out->sweep_block = std::max(1, b->cfg.sweep_block);
    out->sweep_block_mom = b->cfg.sweep_block_mom;
    for (int q = 0; q < 4; ++q) out->n_iter[q] = b->cfg.n_iter[q];
    out->p_solver = IFL_PSOLVE_SWEEPS;
    if (b->cfg.dim != 2 || b->cfg.smoother != IFL_SMOOTHER_RBGS) return IFL_OK;   // nothing to choose
    int st;
    for (ifl_solver *h : b->cases) IFL_HIP(hipStreamSynchronize(h->stream));
    std::vector<Snapshot> snaps(n);
    for (int i = 0; i < n; ++i)
        if ((st = snapTake(b->cases[i], snaps[i])) != IFL_OK) return st;
    struct Restore {
        ifl_batch *b;
        std::vector<ifl_config> c0;
        std::vector<int> t0;
        ifl_config bc0;
        int bt0;
        bool armed = true;
        ~Restore() {
            if (!armed) return;
            for (size_t i = 0; i < c0.size(); ++i) {
                b->cases[i]->cfg = c0[i];
                b->cases[i]->tile = t0[i];
                b->cases[i]->graphDirty = true;
            }
            b->cfg = bc0;
            b->tile = bt0;
        }
    } guard{b, {}, {}, b->cfg, b->tile};
    for (ifl_solver *h : b->cases) {
        guard.c0.push_back(h->cfg);
        guard.t0.push_back(h->tile);
    }
    auto apply = [&](int tile, int K, int Km) {
        for (ifl_solver *h : b->cases) {
            h->cfg.tile = tile;
            h->tile = tile;
            h->cfg.sweep_block = K;
            h->cfg.sweep_block_mom = Km;
            h->graphDirty = true;
        }
        b->cfg.tile = tile;
        b->cfg.sweep_block = K;
        b->cfg.sweep_block_mom = Km;
        b->tile = tile;
    };
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) all[i] = i;
    auto timeOne = [&](double &us) -> int {
        int r;
        for (ifl_solver *h : b->cases)
            if ((r = armCase(h, b->stream, 0)) != IFL_OK) return r;
        if ((r = batchBuild(b, all)) != IFL_OK) return r;
        for (ifl_solver *h : b->cases)
            if (h->iter == 0) {
                Plan pl = planOf(h);
                pl.L.s = b->stream;
                ifl::launch_init_faces(pl.L);
            }
        for (int q = 0; q < 3; ++q) IFL_HIP(hipGraphLaunch(b->graph, b->stream));
        IFL_HIP(hipEventRecord(b->ev0, b->stream));
        for (int q = 0; q < trial; ++q) IFL_HIP(hipGraphLaunch(b->graph, b->stream));
        IFL_HIP(hipEventRecord(b->ev1, b->stream));
        IFL_HIP(hipEventSynchronize(b->ev1));
        float ms = 0;
        IFL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        us = 1000.0 * ms / trial;
        for (int i = 0; i < n; ++i)
            if ((r = snapRestore(b->cases[i], snaps[i])) != IFL_OK) return r;
        out->trials++;
        return IFL_OK;
    };
    double bestUs = 1e300;
    int bt = b->tile, bK = out->sweep_block, bKm = out->sweep_block_mom;
    for (int tile : {ifl::TILE_128x32, ifl::TILE_64x16, ifl::TILE_64x8, ifl::TILE_32x8, ifl::TILE_32x4})
        for (int K : {2, 4, 5})
            for (int Km : {0, 4, 5}) {
                if (Km > K) continue;
                apply(tile, K, Km);
                double us;
                if ((st = timeOne(us)) != IFL_OK) return st;
                if (us < bestUs) {
                    bestUs = us;
                    bt = tile;
                    bK = K;
                    bKm = Km;
                }
            }
    apply(bt, bK, bKm);
    guard.armed = false;
    if (b->graph) {
        IFL_HIP(hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
    }
    out->tile = bt;
    out->sweep_block = bK;
    out->sweep_block_mom = bKm;
    out->us_per_iter = bestUs;
    return IFL_OK;
}

}  // extern "C"
