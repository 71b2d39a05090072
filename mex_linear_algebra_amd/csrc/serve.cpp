// Serving loop: continuous batching of mixed rank / null-space / matmul
// requests under an HBM budget (no reference counterpart; the reference
// serves one synchronous MEX call at a time).
//
// submit() admits nothing by itself: it queues the request (rejecting one
// that could never fit the budget).  A scheduler thread admits queued
// requests in arrival order while their device-memory estimate fits the
// budget, groups the admitted rank and null requests into micro-batches for
// the persistent batch kernel (up to max_batch each), and hands jobs to
// worker threads.  Each worker owns a context (its own HIP stream), so jobs
// run concurrently on the device; requests that arrive while others run are
// admitted as memory and workers free up.  Results stay on the device until
// the client releases them.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "common.h"

using namespace mela;
using Clock = std::chrono::steady_clock;

namespace {

constexpr size_t kBatchMaxM = 4096, kBatchMaxN = 8192;

struct Req {
    uint64_t id = 0;
    mela_request r{};
    size_t est = 0;        // admission estimate (bytes)
    size_t out_bytes = 0;  // bytes of the result kept until release
    Clock::time_point t_submit;
    enum State { PENDING, ADMITTED, DONE } state = PENDING;
    bool cancelled = false;
    mela_result res{};
    std::string err;
};

struct Job {
    enum Kind { BATCH_RANK, BATCH_NULL, SINGLE } kind;
    std::vector<uint64_t> ids;
};

size_t est_bytes(const mela_request &r) {
    const size_t m = r.A.rows, n = r.A.cols, wc = words64(n) ? words64(n) : 1;
    if (r.op == MELA_OP_MATMUL) {
        const size_t nn = r.B.cols, k = r.B.rows;
        return m * (words64(nn) ? words64(nn) : 1) * 8 + nn * (words64(k) ? words64(k) : 1) * 8 + 4096;
    }
    size_t b = m * wc * 8 + std::min(m, n) * wc * 8 + 4096;
    if (r.op == MELA_OP_NULL) b += n * wc * 8;
    return b;
}

// Time (ms) that one workgroup of the persistent batch kernel takes over
// one matrix, end to end (M4RI mode), fitted to single-matrix batches on
// MI355X over 128..4096 rows x 128..8192 columns (within 35 %): a fixed
// cost, a cost per 64-pivot panel, per panel and row (the selection scans
// the active rows), per row, pivot and column word (the sweep) and per word
// of the null basis written.
double batch_member_ms(const mela_request &r) {
    const double m = (double)r.A.rows, n = (double)r.A.cols;
    const double rk = std::min(m, n), wc = (double)words64(r.A.cols);
    const double p = std::ceil(rk / 64.0);
    if (r.op == MELA_OP_NULL)
        return 0.098 + 0.0139 * p + 0.0142e-3 * m * p + 0.0201e-6 * m * rk * wc + 1.514e-6 * n * wc;
    return 0.113 + 0.0054 * p + 0.0155e-3 * m * p + 0.0117e-6 * m * rk * wc + 0.1005e-6 * n * wc;
}

// A batch launch lasts as long as its slowest member, so a matrix joins a
// micro-batch only while one workgroup finishes it about as fast as the
// single-matrix path, which spreads each panel over the whole device
// (0.06-0.4 ms for these small shapes, 1.4-2 ms at 4096 x 8192 where one
// workgroup would take 28-52 ms).
constexpr double kBatchMemberMaxMs = 0.35;

bool batchable(const mela_request &r) {
    return r.op != MELA_OP_MATMUL && r.A.rows <= kBatchMaxM && r.A.cols <= kBatchMaxN &&
           batch_member_ms(r) <= kBatchMemberMaxMs;
}

}  // namespace

struct mela_server {
    int device = 0;
    size_t budget = 0;
    int max_batch = 64;
    std::mutex mu;
    std::condition_variable cv_sched, cv_jobs, cv_done;
    std::deque<uint64_t> pending;
    std::deque<Job> jobs;
    std::map<uint64_t, Req> reqs;
    uint64_t next_id = 1;
    size_t in_use = 0, peak = 0;
    bool stop = false;
    int busy_workers = 0;
    int nworkers = 0;
    std::thread sched;
    std::vector<std::thread> workers;
    // stats
    uint64_t n_submitted = 0, n_completed = 0, n_rejected = 0, n_cancelled = 0, n_failed = 0;
    std::vector<double> lat_ms;
    uint64_t n_jobs = 0, n_batched_reqs = 0;
};

namespace {

void finish(mela_server *s, Req &q, mela_status st, const std::string &err) {
    // caller holds s->mu
    q.state = Req::DONE;
    q.res.status = st;
    q.err = err;
    q.res.latency_ms =
        std::chrono::duration<double, std::milli>(Clock::now() - q.t_submit).count();
    s->in_use -= (q.est - (st == MELA_OK ? q.out_bytes : 0));
    if (st != MELA_OK) q.out_bytes = 0;
    if (q.cancelled && q.res.out.data) {  // nobody will read it
        (void)hipFree(q.res.out.data);
        q.res.out.data = nullptr;
        s->in_use -= q.out_bytes;
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

void scheduler_main(mela_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        // admit in arrival order while the estimate fits the budget
        std::vector<uint64_t> admitted;
        while (!s->pending.empty()) {
            Req &q = s->reqs[s->pending.front()];
            if (s->in_use + q.est > s->budget) break;
            s->pending.pop_front();
            q.state = Req::ADMITTED;
            s->in_use += q.est;
            s->peak = std::max(s->peak, s->in_use);
            admitted.push_back(q.id);
        }
        if (!admitted.empty()) {
            Job br{Job::BATCH_RANK, {}}, bn{Job::BATCH_NULL, {}};
            for (uint64_t id : admitted) {
                const mela_request &r = s->reqs[id].r;
                if (batchable(r)) {
                    Job &j = r.op == MELA_OP_NULL ? bn : br;
                    j.ids.push_back(id);
                    if ((int)j.ids.size() >= s->max_batch) {
                        s->jobs.push_back(j);
                        j.ids.clear();
                    }
                } else {
                    s->jobs.push_back(Job{Job::SINGLE, {id}});
                }
            }
            if (!br.ids.empty()) s->jobs.push_back(br);
            if (!bn.ids.empty()) s->jobs.push_back(bn);
            s->cv_jobs.notify_all();
        }
        s->cv_sched.wait(lk);
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
        q.out_bytes = Z.rows * Z.ld * 8;
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
    q.out_bytes = C.rows * C.ld * 8;
    return MELA_OK;
}

void worker_main(mela_server *s) {
    mela_ctx *ctx = nullptr;
    mela_status cst = mela_ctx_create(s->device, &ctx);
    std::string cerr = cst == MELA_OK ? "" : mela_last_error();
    if (ctx) ctx->capture_ok = false;  // other workers issue synchronous HIP calls
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_jobs.wait(lk, [&] { return s->stop || !s->jobs.empty(); });
        if (s->jobs.empty() && s->stop) break;
        Job job = s->jobs.front();
        s->jobs.pop_front();
        if (job.kind != Job::SINGLE) {
            // continuous batching: fold the micro-batches of the same kind
            // that queued up while every worker was busy into this launch
            for (auto it = s->jobs.begin(); it != s->jobs.end();) {
                if (it->kind == job.kind && job.ids.size() + it->ids.size() <= (size_t)s->max_batch) {
                    job.ids.insert(job.ids.end(), it->ids.begin(), it->ids.end());
                    it = s->jobs.erase(it);
                } else {
                    ++it;
                }
            }
        }
        ++s->busy_workers;
        ++s->n_jobs;
        // snapshot the requests (the map entries stay put while admitted)
        std::vector<Req *> rq;
        for (uint64_t id : job.ids) rq.push_back(&s->reqs[id]);
        lk.unlock();

        std::vector<mela_status> sts(rq.size(), MELA_OK);
        std::vector<std::string> errs(rq.size());
        if (cst != MELA_OK) {
            for (size_t i = 0; i < rq.size(); ++i) {
                sts[i] = cst;
                errs[i] = cerr;
            }
        } else if (job.kind == Job::SINGLE) {
            sts[0] = run_single(ctx, *rq[0]);
            if (sts[0] != MELA_OK) errs[0] = mela_last_error();
        } else {
            std::vector<mela_bitmat> mats;
            for (Req *q : rq) mats.push_back(q->r.A);
            const bool want_null = job.kind == Job::BATCH_NULL;
            mela_batch *b = nullptr;
            mela_status st = mela_batch_create(
                ctx, mats.data(), mats.size(),
                (want_null ? MELA_BATCH_NULL : 0) | MELA_BATCH_CTX_SCRATCH, &b);
            if (st == MELA_OK) st = mela_batch_run(b);
            std::vector<uint32_t> ranks(rq.size());
            std::vector<mela_bitmat> nulls(rq.size());
            if (st == MELA_OK) st = mela_batch_results(b, ranks.data(), nulls.data());
            // copy each null basis out of the batch arena (issued together,
            // one synchronisation for the whole batch)
            std::vector<mela_bitmat> outs(rq.size());
            for (size_t i = 0; i < rq.size() && st == MELA_OK; ++i) {
                rq[i]->res.rank = ranks[i];
                if (!want_null) continue;
                mela_status s2 = mela_bitmat_alloc(ctx, nulls[i].rows, nulls[i].cols, &outs[i]);
                const mela_bitmat &Z = outs[i];
                if (s2 == MELA_OK && Z.rows && nulls[i].cols) {
                    // the batch view has the row pitch of the input width
                    hipError_t e = hipMemcpy2DAsync(
                        Z.data, Z.ld * 8, nulls[i].data, nulls[i].ld * 8, Z.ld * 8, Z.rows,
                        hipMemcpyDeviceToDevice, (hipStream_t)mela_ctx_stream(ctx));
                    if (e != hipSuccess) s2 = hip_fail(e, "hipMemcpy2DAsync");
                }
                if (s2 != MELA_OK) {
                    sts[i] = s2;
                    errs[i] = mela_last_error();
                }
            }
            if (st == MELA_OK && want_null) {
                mela_status s3 = mela_ctx_sync(ctx);
                const std::string e3 = s3 == MELA_OK ? "" : mela_last_error();
                for (size_t i = 0; i < rq.size(); ++i) {
                    if (s3 != MELA_OK && sts[i] == MELA_OK) {
                        sts[i] = s3;
                        errs[i] = e3;
                    }
                    if (sts[i] != MELA_OK) {
                        if (outs[i].data) mela_bitmat_free(ctx, &outs[i]);
                        continue;
                    }
                    rq[i]->res.out = outs[i];
                    rq[i]->out_bytes = outs[i].rows * outs[i].ld * 8;
                }
            }
            if (st != MELA_OK)
                for (size_t i = 0; i < rq.size(); ++i) {
                    sts[i] = st;
                    errs[i] = mela_last_error();
                }
            if (b) mela_batch_destroy(b);
        }

        lk.lock();
        --s->busy_workers;
        if (job.kind != Job::SINGLE) s->n_batched_reqs += rq.size();
        for (size_t i = 0; i < rq.size(); ++i) finish(s, *rq[i], sts[i], errs[i]);
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
    s->sched = std::thread(scheduler_main, s);
    for (int i = 0; i < workers; ++i) s->workers.emplace_back(worker_main, s);
    *out = s;
    return MELA_OK;
}

mela_status mela_server_submit(mela_server *s, const mela_request *req, uint64_t *ticket) {
    if (!s || !req || !ticket) return set_error(MELA_ERR_INVALID, "null argument");
    const mela_request &r = *req;
    if (r.op < MELA_OP_RANK || r.op > MELA_OP_MATMUL)
        return set_error(MELA_ERR_INVALID, "unknown request op");
    if (r.op == MELA_OP_MATMUL && r.A.cols != r.B.rows)
        return set_error(MELA_ERR_DIM, "Inner dimensions must agree.");
    Req q;
    q.r = r;
    q.est = est_bytes(r);
    q.t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return set_error(MELA_ERR_CANCELLED, "server is shutting down");
    if (q.est > s->budget) {
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
    mela_bitmat out{};
    size_t bytes = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return set_error(MELA_ERR_INVALID, "unknown ticket");
        if (it->second.state != Req::DONE) return set_error(MELA_ERR_INVALID, "request still running");
        out = it->second.res.out;
        bytes = it->second.out_bytes;
        s->reqs.erase(it);
    }
    if (out.data) {
        (void)hipSetDevice(s->device);
        (void)hipFree(out.data);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->in_use -= bytes;
    s->cv_sched.notify_all();
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
    st->hbm_in_use = s->in_use;
    st->hbm_peak = s->peak;
    st->jobs = s->n_jobs;
    st->batched_requests = s->n_batched_reqs;
    return MELA_OK;
}

mela_status mela_server_reset_stats(mela_server *s) {
    if (!s) return set_error(MELA_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->n_submitted = s->n_completed = s->n_rejected = s->n_cancelled = s->n_failed = 0;
    s->lat_ms.clear();
    s->peak = s->in_use;
    s->n_jobs = s->n_batched_reqs = 0;
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
    (void)hipSetDevice(s->device);
    for (auto &kv : s->reqs)
        if (kv.second.res.out.data) (void)hipFree(kv.second.res.out.data);
    delete s;
}

}  // extern "C"
