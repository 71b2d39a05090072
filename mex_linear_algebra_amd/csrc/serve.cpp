// Serving: continuous batching of mixed rank / null-space / matmul requests
// under an HBM budget (no reference counterpart; the reference serves one
// synchronous MEX call at a time, e.g. mela_rank_m4ri.c:286,
// mela_null_gf2.c:142, mela_matmul_gf2.c:72).
//
// submit() queues a request (rejecting one that could never fit the budget).
// A scheduler thread admits queued requests in arrival order while the
// device memory they need fits the budget, and routes each one:
//  * rank and null requests that one workgroup finishes quickly go to the
//    device-resident loop: a persistent kernel (batch.hip, k_rref_loop with the serving source)
//    whose workgroups take request slots from a ring in pinned host memory
//    as the scheduler publishes them, so requests that arrive while others
//    are on the GPU start as soon as a workgroup is free, with no launch per
//    request or per batch.  A poller thread sees the completions and starts
//    the loop again when new work arrives after it went idle.
//  * larger rank / null requests and all products go to worker threads, each
//    with its own context and stream (the single-matrix entry points, which
//    spread one request over the whole device).
// Accounting (all device bytes the server holds, which admission keeps at or
// below the budget): the loop's buffer pool (blocks in use and cached), each
// worker context's scratch cache, outputs waiting for release, and for every
// running worker job an upper bound of the scratch and output it may add.
// Device memory is freed outside the server lock.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "common.h"

using namespace mela;
using Clock = std::chrono::steady_clock;

namespace {

constexpr size_t kLoopMaxM = 4096, kLoopMaxN = 8192;
constexpr uint32_t kRingCap = 4096;           // request slots of the device loop
constexpr uint64_t kLoopIdleTicks = 50000;   // 0.5 ms (100 MHz clock) without work ends the loop
constexpr uint64_t kLoopLifeTicks = 100000;   // a launch takes no new request after 1 ms
constexpr size_t kPoolMinBlock = (size_t)64 << 10;

size_t pool_class(size_t bytes) {
    size_t c = kPoolMinBlock;
    while (c < bytes) c <<= 1;
    return c;
}

size_t ld16(size_t n) { return (words64(n) + 15) & ~(size_t)15; }

// Upper bound of the device bytes a worker job adds while it runs: scratch
// the library may grow (working copy, pivot rows, metadata and active list
// of an elimination, the blocked path's block state and operands; re-laid,
// expanded or transposed operands of a product) plus the output it
// allocates.  A scratch slot that grows at least doubles, so it can end at
// up to twice what the job needs.
size_t est_single(const mela_request &r) {
    const size_t m = r.A.rows, n = r.A.cols;
    if (r.op == MELA_OP_MATMUL) {
        const size_t k = r.A.cols, nn = r.B.cols;
        const size_t mp = (m + 255) / 256 * 256, np = (nn + 255) / 256 * 256;
        const size_t out = std::max<size_t>(m, 1) * std::max<size_t>(words64(nn), 1) * 8;
        const size_t tiled = (mp + np) * ((k + 1023) / 1024 * 1024) / 8;      // fp4, packed planes
        const size_t expanded = (mp + np) * ((k + 511) / 512 * 512) / 2;      // fp4, small grids
        const size_t bt = nn * std::max<size_t>(words64(k), 1) * 8;           // AND+parity
        return out + 2 * (std::max(tiled, std::max(expanded, bt)) + ((size_t)1 << 20));
    }
    const size_t rmax = std::min(m, n);
    const size_t rows = m + rmax;
    size_t scratch = rows * ld16(n) * 8 + m * 4 + rmax * 4 + ((size_t)64 << 10);
    scratch += rows * 64 + n * 64 + ((size_t)64 << 10);  // blocked path: operands, block state
    size_t b = 2 * scratch;
    if (r.op == MELA_OP_NULL) b += n * std::max<size_t>(words64(n), 1) * 8 + n * 8 + words64(n) * 80;
    return b;
}

// Device-loop buffers of a rank / null request: the working rows and pivot
// rows, and the null basis (n rows of ceil(n/64) words, the widest it can be).
size_t loop_work_bytes(const mela_request &r) {
    const size_t m = r.A.rows, n = r.A.cols, wc = std::max<size_t>(words64(n), 1);
    return (m + std::min(m, n)) * wc * 8 + 8;
}
size_t loop_z_bytes(const mela_request &r) {
    return r.op == MELA_OP_NULL ? r.A.cols * std::max<size_t>(words64(r.A.cols), 1) * 8 + 8 : 0;
}

// Time (ms) that one workgroup of the batch machinery takes over one matrix,
// end to end (M4RI mode), fitted to single-matrix batches on MI355X over
// 128..4096 rows x 128..8192 columns (within 35 %): a fixed cost, a cost per
// 64-pivot panel, per panel and row (the selection scans the active rows),
// per row, pivot and column word (the sweep) and per word of the null basis.
double loop_member_ms(const mela_request &r) {
    const double m = (double)r.A.rows, n = (double)r.A.cols;
    const double rk = std::min(m, n), wc = (double)words64(r.A.cols);
    const double p = std::ceil(rk / 64.0);
    if (r.op == MELA_OP_NULL)
        return 0.098 + 0.0139 * p + 0.0142e-3 * m * p + 0.0201e-6 * m * rk * wc + 1.514e-6 * n * wc;
    return 0.113 + 0.0054 * p + 0.0155e-3 * m * p + 0.0117e-6 * m * rk * wc + 0.1005e-6 * n * wc;
}

// A request goes to the device loop while one workgroup finishes it about
// as fast as the single-matrix path, which spreads each panel over the whole
// device (0.06-0.4 ms for small shapes, 1.4-2 ms at 4096 x 8192 where one
// workgroup would take 28-52 ms).
constexpr double kLoopMemberMaxMs = 0.35;

bool loop_eligible(const mela_request &r) {
    return r.op != MELA_OP_MATMUL && r.A.rows && r.A.cols && r.A.rows <= kLoopMaxM &&
           r.A.cols <= kLoopMaxN && loop_member_ms(r) <= kLoopMemberMaxMs;
}

struct Req {
    uint64_t id = 0;
    mela_request r{};
    bool loop = false;      // device loop (else a worker job)
    size_t reserve = 0;     // worker job: bytes reserved while it runs
    size_t out_bytes = 0;   // worker output held until release
    void *work = nullptr;   // loop: pool blocks (work rows, null basis)
    void *z = nullptr;
    size_t work_cls = 0, z_cls = 0;
    uint64_t seq = 0;       // loop slot sequence number
    Clock::time_point t_submit;
    enum State { PENDING, ADMITTED, DONE } state = PENDING;
    bool cancelled = false;
    mela_result res{};
    std::string err;
};

}  // namespace

struct mela_server {
    int device = 0;
    size_t budget = 0;
    int max_batch = 64;
    std::mutex mu;
    std::condition_variable cv_sched, cv_jobs, cv_done;
    std::deque<uint64_t> pending;
    std::deque<uint64_t> jobs;  // admitted worker jobs
    std::map<uint64_t, Req> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread sched, poller;
    std::vector<std::thread> workers;
    int nworkers = 0;
    // accounting (bytes)
    size_t pool_held = 0;       // every pool block, in use or cached
    size_t reserved = 0;        // running worker jobs' upper bounds
    size_t outputs = 0;         // worker outputs waiting for release
    std::vector<size_t> worker_cache;  // scratch held by each worker context
    int warm_left = 0;                 // workers still warming up (mela_server_create waits)
    std::vector<char> worker_idle, worker_drop;
    size_t peak = 0;
    std::map<size_t, std::vector<void *>> pool_free;  // cached blocks by size class
    // device loop
    mela_ctx *loop_ctx = nullptr;  // stream the loop kernel runs on
    std::string loop_err;
    ServeSlot *slots = nullptr;
    ServeDone *done = nullptr;
    uint64_t *tail = nullptr;    // pinned host: slots published
    uint64_t *d_head = nullptr;  // device: slots taken
    uint64_t published = 0;
    std::vector<uint64_t> inflight;  // ids of loop requests on the device
    hipEvent_t loop_ev = nullptr;
    bool loop_launched = false;
    int loop_grid = 256;
    // stats
    uint64_t n_submitted = 0, n_completed = 0, n_rejected = 0, n_cancelled = 0, n_failed = 0;
    std::vector<double> lat_ms;
    uint64_t n_jobs = 0, n_loop_reqs = 0, n_loop_launches = 0;

    size_t held() const {
        size_t h = pool_held + reserved + outputs;
        for (size_t c : worker_cache) h += c;
        return h;
    }
    void note_peak() { peak = std::max(peak, held()); }
};

namespace {

// caller holds s->mu; device pointers to free once it is released
void finish(mela_server *s, Req &q, mela_status st, const std::string &err,
            std::vector<void *> &to_free) {
    q.state = Req::DONE;
    q.res.status = st;
    q.err = err;
    q.res.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - q.t_submit).count();
    if (q.cancelled || st != MELA_OK) {
        // nobody will read the output
        if (q.loop && q.z) {
            s->pool_free[q.z_cls].push_back(q.z);
            q.z = nullptr;
        } else if (!q.loop && q.res.out.data) {
            to_free.push_back(q.res.out.data);
            s->outputs -= q.out_bytes;
        }
        q.res.out.data = nullptr;
        q.out_bytes = 0;
    }
    if (q.cancelled) {
        ++s->n_cancelled;
    } else if (st == MELA_OK) {
        ++s->n_completed;
        s->lat_ms.push_back(q.res.latency_ms);
    } else {
        ++s->n_failed;
    }
    s->cv_done.notify_all();
    s->cv_sched.notify_all();
}

// Makes `device` current on the calling thread.  A server created on a
// device that does not exist still runs (its requests fail); the failed
// call's error is cleared so that it does not surface in the caller's next,
// unrelated HIP call.
void use_device(int device) {
    if (hipSetDevice(device) != hipSuccess) (void)hipGetLastError();
}

void free_all(const std::vector<void *> &ptrs, int device) {
    if (ptrs.empty()) return;
    use_device(device);
    for (void *p : ptrs) (void)hipFree(p);
}

// caller holds s->mu: starts the loop kernel when requests are on the
// device and no instance is running (a finished instance's event is
// complete).  A workgroup ends only while it holds no request, so requests
// still unreported after the kernel ended were never taken.
void ensure_loop(mela_server *s) {
    if (!s->loop_ctx || s->inflight.empty()) return;
    if (s->loop_launched) {
        const hipError_t q = hipEventQuery(s->loop_ev);
        if (q == hipErrorNotReady) return;  // still running: it will take the slots
        (void)hipGetLastError();
    }
    use_device(s->device);
    hipStream_t st = (hipStream_t)mela_ctx_stream(s->loop_ctx);
    if (launch_serve_loop(st, s->loop_grid, s->slots, s->done, s->tail, s->d_head, kRingCap, kLoopIdleTicks,
                          kLoopLifeTicks) != MELA_OK ||
        hipEventRecord(s->loop_ev, st) != hipSuccess) {
        (void)hipGetLastError();
        s->loop_err = mela_last_error();
        return;
    }
    s->loop_launched = true;
    ++s->n_loop_launches;
}

// caller holds lk on s->mu.  Blocks of sizes cw (and cz) from the pool, or
// fresh ones if the budget allows their growth; false if neither (nothing
// taken).  hipMalloc runs with the lock released.
bool pool_take(mela_server *s, std::unique_lock<std::mutex> &lk, size_t cw, size_t cz, void **pw,
               void **pz) {
    auto &fw = s->pool_free[cw];
    size_t grow;
    if (cz == cw) {  // both blocks of one class
        const size_t have = fw.size();
        grow = (have >= 2 ? 0 : 2 - have) * cw;
    } else {
        grow = (fw.empty() ? cw : 0) + (cz && s->pool_free[cz].empty() ? cz : 0);
    }
    if (s->held() + grow > s->budget) return false;
    *pw = *pz = nullptr;
    if (!fw.empty()) {
        *pw = fw.back();
        fw.pop_back();
    }
    if (cz) {
        auto &fz = s->pool_free[cz];
        if (!fz.empty()) {
            *pz = fz.back();
            fz.pop_back();
        }
    }
    const size_t need = (*pw ? 0 : cw) + (cz && !*pz ? cz : 0);
    if (need == 0) return true;
    s->pool_held += need;  // reserved before the allocation
    s->note_peak();
    lk.unlock();
    use_device(s->device);
    bool ok = true;
    if (!*pw && hipMalloc(pw, cw) != hipSuccess) {
        *pw = nullptr;
        ok = false;
    }
    if (ok && cz && !*pz && hipMalloc(pz, cz) != hipSuccess) {
        *pz = nullptr;
        ok = false;
    }
    if (!ok) (void)hipGetLastError();
    lk.lock();
    if (!ok) {
        // give back what was reserved; keep what was allocated in the cache
        s->pool_held -= need;
        if (*pw) {
            s->pool_held += cw;
            s->pool_free[cw].push_back(*pw);
        }
        if (*pz) {
            s->pool_held += cz;
            s->pool_free[cz].push_back(*pz);
        }
        *pw = *pz = nullptr;
    }
    return ok;
}

// caller holds lk.  Frees cached pool blocks (outside the lock) and asks
// idle workers to drop their scratch caches; true if anything was freed.
bool reclaim(mela_server *s, std::unique_lock<std::mutex> &lk) {
    std::vector<void *> blocks;
    for (auto &kv : s->pool_free) {
        for (void *p : kv.second) {
            blocks.push_back(p);
            s->pool_held -= kv.first;
        }
        kv.second.clear();
    }
    bool asked = false;
    for (int w = 0; w < s->nworkers; ++w)
        if (s->worker_cache[(size_t)w] && s->worker_idle[(size_t)w] && !s->worker_drop[(size_t)w]) {
            s->worker_drop[(size_t)w] = 1;
            asked = true;
        }
    if (asked) s->cv_jobs.notify_all();
    if (blocks.empty()) return false;
    lk.unlock();
    free_all(blocks, s->device);
    lk.lock();
    return true;
}

// caller holds s->mu: a slot may be written once the request that used it
// kRingCap sequence numbers earlier has been reported
bool ring_has_room(const mela_server *s) {
    if (s->inflight.size() >= (size_t)std::min<int>(s->max_batch, (int)kRingCap)) return false;
    uint64_t oldest = s->published;
    for (uint64_t id : s->inflight) oldest = std::min(oldest, s->reqs.at(id).seq);
    return s->published - oldest < kRingCap;
}

void scheduler_main(mela_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        while (!s->pending.empty() && !s->stop) {
            const uint64_t id = s->pending.front();
            Req &q = s->reqs[id];
            if (q.loop && !s->loop_ctx) {
                // no device for the loop: the request fails with its error
                s->pending.pop_front();
                q.state = Req::ADMITTED;
                ++s->n_loop_reqs;
                std::vector<void *> none;
                finish(s, q, MELA_ERR_DEVICE, s->loop_err, none);
                continue;
            }
            if (q.loop) {
                if (!ring_has_room(s)) break;  // the poller's completions wake us
                const size_t cw = pool_class(loop_work_bytes(q.r));
                const size_t cz = q.r.op == MELA_OP_NULL ? pool_class(loop_z_bytes(q.r)) : 0;
                void *pw = nullptr, *pz = nullptr;
                if (!pool_take(s, lk, cw, cz, &pw, &pz)) {
                    if (reclaim(s, lk)) continue;  // freed cached blocks: try again
                    break;  // wait for completions and releases
                }
                if (s->reqs.count(id) == 0 || s->pending.empty() || s->pending.front() != id) {
                    // cancelled while the lock was released: the blocks go back
                    if (pw) s->pool_free[cw].push_back(pw);
                    if (pz) s->pool_free[cz].push_back(pz);
                    continue;
                }
                s->pending.pop_front();
                q.work = pw;
                q.z = pz;
                q.work_cls = cw;
                q.z_cls = cz;
                q.state = Req::ADMITTED;
                q.seq = s->published;
                ServeSlot &sl = s->slots[q.seq % kRingCap];
                sl.A = q.r.A.data;
                sl.work = (uint64_t *)pw;
                sl.Z = (uint64_t *)pz;
                sl.m = (uint32_t)q.r.A.rows;
                sl.n = (uint32_t)q.r.A.cols;
                sl.lda = (uint32_t)q.r.A.ld;
                sl.op = (uint32_t)q.r.op;
                sl.mode = q.r.variant == MELA_ELIM_GE ? 0u : 1u;
                s->done[q.seq % kRingCap].seq1 = 0;
                ++s->published;
                __atomic_store_n(s->tail, s->published, __ATOMIC_RELEASE);
                s->inflight.push_back(id);
                ++s->n_loop_reqs;
                ensure_loop(s);
            } else {
                const size_t est = est_single(q.r);
                if (s->held() + est > s->budget) {
                    if (reclaim(s, lk)) continue;
                    break;
                }
                s->pending.pop_front();
                q.state = Req::ADMITTED;
                q.reserve = est;
                s->reserved += est;
                s->jobs.push_back(id);
                s->cv_jobs.notify_one();
            }
            s->note_peak();
        }
        s->cv_sched.wait(lk);
    }
}

// completions of the device loop, and restarts of the loop kernel
void poller_main(mela_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        if (s->stop && s->inflight.empty()) break;
        std::vector<void *> none;
        bool any = false;
        for (size_t i = 0; i < s->inflight.size();) {
            Req &q = s->reqs[s->inflight[i]];
            const uint64_t seq1 = __atomic_load_n(&s->done[q.seq % kRingCap].seq1, __ATOMIC_ACQUIRE);
            if (seq1 != q.seq + 1) {
                ++i;
                continue;
            }
            q.res.rank = s->done[q.seq % kRingCap].rank;
            if (q.r.op == MELA_OP_NULL) {
                q.res.out.data = (uint64_t *)q.z;
                q.res.out.rows = q.r.A.cols;
                q.res.out.cols = q.r.A.cols - q.res.rank;
                q.res.out.ld = std::max<size_t>(words64(q.r.A.cols), 1);
            }
            s->pool_free[q.work_cls].push_back(q.work);  // the working rows are cached for reuse
            q.work = nullptr;
            finish(s, q, MELA_OK, "", none);
            s->inflight[i] = s->inflight.back();
            s->inflight.pop_back();
            any = true;
        }
        if (!s->inflight.empty()) ensure_loop(s);
        if (any) s->cv_sched.notify_all();  // ring slots and pool blocks came free
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(s->inflight.empty() ? 200 : 10));
        lk.lock();
    }
}

mela_status run_single(mela_ctx *ctx, Req &q) {
    const mela_request &r = q.r;
    if (r.op == MELA_OP_RANK) {
        size_t rk = 0;
        mela_status st = mela_rank_packed(ctx, &r.A, r.variant, &rk);
        q.res.rank = rk;
        return st;
    }
    if (r.op == MELA_OP_NULL) {
        mela_bitmat Z{};
        size_t f = 0;
        mela_status st = mela_null_packed(ctx, &r.A, r.variant, &Z, &f);
        if (st != MELA_OK) return st;
        q.res.out = Z;
        q.res.rank = r.A.cols - f;
        q.out_bytes = std::max<size_t>(Z.rows, 1) * Z.ld * 8;
        return MELA_OK;
    }
    mela_bitmat C{};
    mela_status st = mela_bitmat_alloc(ctx, r.A.rows, r.B.cols, &C);
    if (st != MELA_OK) return st;
    st = mela_matmul_packed(ctx, &r.A, &r.B, &C, r.variant);
    if (st == MELA_OK) st = mela_ctx_sync(ctx);
    if (st != MELA_OK) {
        mela_bitmat_free(ctx, &C);
        return st;
    }
    q.res.out = C;
    q.out_bytes = std::max<size_t>(C.rows, 1) * C.ld * 8;
    return MELA_OK;
}

// A process's first launch of a kernel loads its code and a context's first
// calls grow its scratch; a worker does both before it takes requests (rank
// and null of a dense 2048 x 2048 matrix on the blocked and on the panel
// path, and its square), so that early requests do not wait for them.
// Failures here are ignored: the requests themselves report any error.
constexpr size_t kWarmBudgetPerWorker = (size_t)256 << 20;

void warm_worker(mela_ctx *ctx) {
    constexpr size_t kN = 2048;
    mela_bitmat A{}, Z{}, C{};
    if (mela_bitmat_alloc(ctx, kN, kN, &A) != MELA_OK) return;
    if (mela_gen_bernoulli(ctx, 7, &A) == MELA_OK) {
        for (int v : {MELA_ELIM_AUTO, MELA_ELIM_PANEL}) {
            size_t r = 0, f = 0;
            (void)mela_rank_packed(ctx, &A, v, &r);
            if (mela_null_packed(ctx, &A, v, &Z, &f) == MELA_OK) mela_bitmat_free(ctx, &Z);
        }
        if (mela_bitmat_alloc(ctx, kN, kN, &C) == MELA_OK) {
            (void)mela_matmul_packed(ctx, &A, &A, &C, MELA_ALGO_AUTO);
            mela_bitmat_free(ctx, &C);
        }
    }
    mela_bitmat_free(ctx, &A);
}

void worker_main(mela_server *s, int w) {
    mela_ctx *ctx = nullptr;
    mela_status cst = mela_ctx_create(s->device, &ctx);
    std::string cerr = cst == MELA_OK ? "" : mela_last_error();
    if (ctx) {
        ctx->capture_ok = false;  // other threads issue synchronous HIP calls
        // the warm-up's scratch (a few MiB) is held from the start, so it
        // runs only where the budget leaves ample room for it
        if (s->budget >= (size_t)s->nworkers * kWarmBudgetPerWorker) warm_worker(ctx);
        ctx->defer_free = true;  // no device-wide wait inside a request
    }
    std::unique_lock<std::mutex> lk(s->mu);
    s->worker_cache[(size_t)w] = ctx ? ctx_scratch_bytes(ctx) : 0;
    if (--s->warm_left == 0) s->cv_sched.notify_all();
    for (;;) {
        s->worker_idle[(size_t)w] = 1;
        s->cv_jobs.wait(lk, [&] { return s->stop || !s->jobs.empty() || s->worker_drop[(size_t)w]; });
        if (s->worker_drop[(size_t)w] && ctx) {
            // the scheduler needs the memory: drop the scratch cache
            s->worker_drop[(size_t)w] = 0;
            lk.unlock();
            ctx_release_scratch(ctx);
            lk.lock();
            s->worker_cache[(size_t)w] = 0;
            s->cv_sched.notify_all();
            continue;
        }
        if (s->jobs.empty()) {
            if (s->stop) break;
            continue;
        }
        const uint64_t id = s->jobs.front();
        s->jobs.pop_front();
        s->worker_idle[(size_t)w] = 0;
        ++s->n_jobs;
        Req &q = s->reqs[id];  // map entries stay put while admitted
        lk.unlock();
        mela_status st = cst;
        std::string err = cerr;
        if (cst == MELA_OK) {
            st = run_single(ctx, q);
            if (st != MELA_OK) err = mela_last_error();
        }
        const size_t cache = ctx ? ctx_scratch_bytes(ctx) : 0;
        lk.lock();
        s->reserved -= q.reserve;
        q.reserve = 0;
        s->worker_cache[(size_t)w] = cache;
        if (st == MELA_OK) s->outputs += q.out_bytes;
        std::vector<void *> to_free;
        finish(s, q, st, err, to_free);
        s->note_peak();
        if (!to_free.empty()) {
            lk.unlock();
            free_all(to_free, s->device);
            lk.lock();
        }
    }
    lk.unlock();
    if (ctx) mela_ctx_destroy(ctx);
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

void loop_shutdown(mela_server *s) {
    if (s->tail) __atomic_store_n(s->tail, s->published | SERVE_STOP, __ATOMIC_RELEASE);
    if (s->loop_launched) (void)hipEventSynchronize(s->loop_ev);
    if (s->loop_ev) (void)hipEventDestroy(s->loop_ev);
    if (s->loop_ctx) mela_ctx_destroy(s->loop_ctx);
    for (void *p : {(void *)s->slots, (void *)s->done, (void *)s->tail})
        if (p) (void)hipHostFree(p);
    if (s->d_head) (void)hipFree(s->d_head);
    s->loop_ctx = nullptr;
    s->slots = nullptr;
    s->done = nullptr;
    s->tail = nullptr;
    s->d_head = nullptr;
}

}  // namespace

extern "C" {

mela_status mela_server_create(int device, size_t hbm_budget, int workers, int max_batch,
                               mela_server **out) {
    if (!out || workers < 1 || max_batch < 1 || hbm_budget == 0)
        return set_error(MELA_ERR_INVALID, "bad server parameters");
    mela_server *s = new mela_server();
    s->device = device;
    s->budget = hbm_budget;
    s->max_batch = max_batch;
    s->nworkers = workers;
    s->worker_cache.assign((size_t)workers, 0);
    s->worker_idle.assign((size_t)workers, 1);
    s->worker_drop.assign((size_t)workers, 0);
    // the device loop: its stream, and the ring in pinned (coherent) host memory
    mela_status st = mela_ctx_create(device, &s->loop_ctx);
    if (st == MELA_OK) {
        s->loop_ctx->capture_ok = false;
        s->loop_grid = serve_loop_grid(s->loop_ctx);
        const unsigned fl = hipHostMallocCoherent | hipHostMallocMapped;
        if (hipHostMalloc((void **)&s->slots, kRingCap * sizeof(ServeSlot), fl) != hipSuccess ||
            hipHostMalloc((void **)&s->done, kRingCap * sizeof(ServeDone), fl) != hipSuccess ||
            hipHostMalloc((void **)&s->tail, 64, fl) != hipSuccess ||
            hipMalloc((void **)&s->d_head, sizeof(uint64_t)) != hipSuccess ||
            hipEventCreateWithFlags(&s->loop_ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            loop_shutdown(s);
            delete s;
            return set_error(MELA_ERR_NOMEM, "pinned memory for the serving loop", "mela:memory");
        }
        std::memset(s->slots, 0, kRingCap * sizeof(ServeSlot));
        std::memset(s->done, 0, kRingCap * sizeof(ServeDone));
        std::memset(s->tail, 0, 64);
        const uint64_t zero = 0;
        if (hipMemcpy(s->d_head, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            loop_shutdown(s);
            delete s;
            return set_error(MELA_ERR_DEVICE, "serving loop set-up", "mela:device");
        }
    } else {
        // no device: every request fails with the context's error
        s->loop_err = mela_last_error();
        s->loop_ctx = nullptr;
    }
    s->sched = std::thread(scheduler_main, s);
    s->poller = std::thread(poller_main, s);
    s->warm_left = workers;
    for (int i = 0; i < workers; ++i) s->workers.emplace_back(worker_main, s, i);
    {
        // requests are taken once every worker is warm
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_sched.wait(lk, [&] { return s->warm_left == 0; });
    }
    *out = s;
    return MELA_OK;
}

mela_status mela_server_submit(mela_server *s, const mela_request *req, uint64_t *ticket) {
    if (!s || !req || !ticket) return set_error(MELA_ERR_INVALID, "null argument");
    const mela_request &r = *req;
    if (r.op < MELA_OP_RANK || r.op > MELA_OP_MATMUL)
        return set_error(MELA_ERR_INVALID, "unknown request op");
    if ((r.A.rows && r.A.cols && !r.A.data) || (r.A.cols && r.A.ld < words64(r.A.cols)))
        return set_error(MELA_ERR_INVALID, "bad input descriptor");
    if (r.op == MELA_OP_MATMUL) {
        if (r.A.cols != r.B.rows) return set_error(MELA_ERR_DIM, "Inner dimensions must agree.");
        if ((r.B.rows && r.B.cols && !r.B.data) || (r.B.cols && r.B.ld < words64(r.B.cols)))
            return set_error(MELA_ERR_INVALID, "bad input descriptor");
    }
    Req q;
    q.r = r;
    q.loop = loop_eligible(r);
    q.t_submit = Clock::now();
    const size_t need = q.loop ? pool_class(loop_work_bytes(r)) + (r.op == MELA_OP_NULL ? pool_class(loop_z_bytes(r)) : 0)
                               : est_single(r);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return set_error(MELA_ERR_CANCELLED, "server is shutting down");
    if (need > s->budget) {
        ++s->n_rejected;
        return set_error(MELA_ERR_ADMISSION, "request needs more device memory than the server budget");
    }
    q.id = s->next_id++;
    *ticket = q.id;
    s->reqs[q.id] = q;
    s->pending.push_back(q.id);
    ++s->n_submitted;
    s->cv_sched.notify_one();
    return MELA_OK;
}

mela_status mela_server_wait(mela_server *s, uint64_t ticket, double timeout_s, mela_result *res) {
    if (!s) return set_error(MELA_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return set_error(MELA_ERR_INVALID, "unknown ticket");
    auto done = [&] { return it->second.state == Req::DONE; };
    if (timeout_s < 0)
        s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
        return set_error(MELA_ERR_TIMEOUT, "request not finished before the deadline");
    Req &q = it->second;
    if (res) *res = q.res;
    if (q.cancelled) return set_error(MELA_ERR_CANCELLED, "request cancelled");
    if (q.res.status != MELA_OK) return set_error(q.res.status, q.err);
    return MELA_OK;
}

mela_status mela_server_cancel(mela_server *s, uint64_t ticket) {
    if (!s) return set_error(MELA_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return set_error(MELA_ERR_INVALID, "unknown ticket");
    Req &q = it->second;
    if (q.state == Req::DONE) return MELA_OK;
    q.cancelled = true;
    if (q.state == Req::PENDING) {
        s->pending.erase(std::find(s->pending.begin(), s->pending.end(), ticket));
        q.state = Req::DONE;
        q.res.status = MELA_ERR_CANCELLED;
        ++s->n_cancelled;
        s->cv_done.notify_all();
        s->cv_sched.notify_all();
    }
    return MELA_OK;
}

mela_status mela_server_release(mela_server *s, uint64_t ticket) {
    if (!s) return set_error(MELA_ERR_INVALID, "null server");
    void *to_free = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return set_error(MELA_ERR_INVALID, "unknown ticket");
        Req &q = it->second;
        if (q.state != Req::DONE) return set_error(MELA_ERR_INVALID, "request still running");
        if (q.loop) {
            if (q.z) s->pool_free[q.z_cls].push_back(q.z);  // cached for the next request
        } else if (q.res.out.data) {
            to_free = q.res.out.data;
            s->outputs -= q.out_bytes;
        }
        s->reqs.erase(it);
        s->cv_sched.notify_all();
    }
    if (to_free) free_all({to_free}, s->device);
    return MELA_OK;
}

mela_status mela_server_stats_get(mela_server *s, mela_server_stats *st) {
    if (!s || !st) return set_error(MELA_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    st->submitted = s->n_submitted;
    st->completed = s->n_completed;
    st->rejected = s->n_rejected;
    st->cancelled = s->n_cancelled;
    st->failed = s->n_failed;
    st->p50_ms = percentile(s->lat_ms, 0.50);
    st->p99_ms = percentile(s->lat_ms, 0.99);
    double sum = 0;
    for (double v : s->lat_ms) sum += v;
    st->mean_ms = s->lat_ms.empty() ? 0.0 : sum / (double)s->lat_ms.size();
    st->hbm_in_use = s->held();
    st->hbm_peak = s->peak;
    st->jobs = s->n_jobs;
    st->batched_requests = s->n_loop_reqs;
    st->loop_launches = s->n_loop_launches;
    return MELA_OK;
}

mela_status mela_server_reset_stats(mela_server *s) {
    if (!s) return set_error(MELA_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->n_submitted = s->n_completed = s->n_rejected = s->n_cancelled = s->n_failed = 0;
    s->lat_ms.clear();
    s->peak = s->held();
    s->n_jobs = s->n_loop_reqs = s->n_loop_launches = 0;
    return MELA_OK;
}

void mela_server_destroy(mela_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        // requests not yet admitted are cancelled; admitted ones finish
        for (uint64_t id : s->pending) {
            Req &q = s->reqs[id];
            q.cancelled = true;
            q.state = Req::DONE;
            q.res.status = MELA_ERR_CANCELLED;
        }
        s->pending.clear();
        s->stop = true;
        s->cv_sched.notify_all();
        s->cv_jobs.notify_all();
        s->cv_done.notify_all();
    }
    s->sched.join();
    for (auto &t : s->workers) t.join();
    s->poller.join();  // returns once the loop's requests are all reported
    loop_shutdown(s);
    use_device(s->device);
    std::vector<void *> blocks;
    for (auto &kv : s->reqs) {
        Req &q = kv.second;
        if (q.loop) {
            if (q.z) blocks.push_back(q.z);
            if (q.work) blocks.push_back(q.work);
        } else if (q.res.out.data) {
            blocks.push_back(q.res.out.data);
        }
    }
    for (auto &kv : s->pool_free)
        for (void *p : kv.second) blocks.push_back(p);
    free_all(blocks, s->device);
    delete s;
}

}  // extern "C"
