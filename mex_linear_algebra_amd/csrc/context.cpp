// Context, errors and device memory of libmela.
#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <cstring>

#include "common.h"

namespace {
thread_local std::string t_err;
thread_local std::string t_err_id;
}  // namespace

namespace mela {

mela_status set_error(mela_status st, const std::string &msg, const char *id) {
    t_err = msg;
    t_err_id = id ? id : "";
    return st;
}

mela_status hip_fail(hipError_t e, const char *what) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
    if (e == hipErrorOutOfMemory) return set_error(MELA_ERR_NOMEM, buf, "mela:memory");
    return set_error(MELA_ERR_DEVICE, buf, "mela:device");
}

mela_status scratch_get(mela_ctx *ctx, int slot, size_t bytes, void **out) {
    if (slot < 0 || slot >= SLOT_COUNT) return set_error(MELA_ERR_INVALID, "bad scratch slot");
    if ((int)ctx->scratch.size() < SLOT_COUNT) ctx->scratch.resize(SLOT_COUNT);
    Scratch &sc = ctx->scratch[slot];
    if (bytes == 0) bytes = 16;
    bytes = (bytes + 255) & ~(size_t)255;
    if (sc.bytes < bytes) {
        // a slot that has to grow at least doubles (exact size if that
        // fails), so calls of growing sizes reallocate a few times, not at
        // every new maximum
        const size_t want = bytes;
        const bool regrow = sc.ptr != nullptr;
        if (regrow) bytes = std::max(bytes, 2 * sc.bytes);
        if (sc.ptr) {
            // captured eliminations may point into it
            for (auto &g : ctx->elim_graphs) {
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
                if (g.graph) (void)hipGraphDestroy(g.graph);
            }
            ctx->elim_graphs.clear();
            if (ctx->defer_free) {
                // queued work may still read it; it stays allocated
                ctx->retired.emplace_back(sc.ptr, sc.bytes);
            } else {
                // the previous block may still be read by queued work
                hipError_t e = hipStreamSynchronize(ctx->stream);
                if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
                (void)hipFree(sc.ptr);
            }
            sc.ptr = nullptr;
            sc.bytes = 0;
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess && regrow && bytes > want) {
            (void)hipGetLastError();
            bytes = want;
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(MELA_ERR_NOMEM, "Out of memory.", "mela:memory");
        }
        sc.ptr = p;
        sc.bytes = bytes;
    }
    *out = sc.ptr;
    return MELA_OK;
}

size_t ctx_scratch_bytes(mela_ctx *ctx) {
    CtxLock lk(ctx->mu);
    size_t b = 0;
    for (const auto &sc : ctx->scratch) b += sc.bytes;
    for (const auto &r : ctx->retired) b += r.second;
    return b;
}

void ctx_release_scratch(mela_ctx *ctx) {
    CtxLock lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &g : ctx->elim_graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    ctx->elim_graphs.clear();
    for (auto &sc : ctx->scratch) {
        if (sc.ptr) (void)hipFree(sc.ptr);
        sc.ptr = nullptr;
        sc.bytes = 0;
    }
    for (auto &r : ctx->retired) (void)hipFree(r.first);
    ctx->retired.clear();
}

void pinned_release(mela_ctx *ctx) {
    if (ctx->pin) {
        if (ctx->pin_is_malloc)
            std::free(ctx->pin);
        else
            (void)hipHostFree(ctx->pin);
    }
    ctx->pin = nullptr;
    ctx->pin_bytes = 0;
    ctx->pin_is_malloc = false;
}

mela_status pinned_get(mela_ctx *ctx, size_t bytes, uint8_t **out) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    if (bytes == 0) bytes = 16;
    if (ctx->pin_bytes < bytes) {
        pinned_release(ctx);
        const size_t want = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
        void *p = nullptr;
        e = ctx->force_pageable ? hipErrorOutOfMemory
                                : hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            // no pinned memory to be had: ordinary host memory still works
            // (the copies then go through the runtime's own staging)
            (void)hipGetLastError();
            p = std::malloc(want);
            if (!p) return set_error(MELA_ERR_NOMEM, "Out of memory.", "mela:memory");
            ctx->pin_is_malloc = true;
        }
        ctx->pin = p;
        ctx->pin_bytes = want;
    }
    *out = (uint8_t *)ctx->pin;
    return MELA_OK;
}

}  // namespace mela

using namespace mela;

extern "C" {

const char *mela_last_error(void) { return t_err.c_str(); }
const char *mela_last_error_id(void) { return t_err_id.c_str(); }
const char *mela_version(void) { return "mela-mi355x 0.1 (gfx950)"; }

mela_status mela_ctx_create(int device, mela_ctx **out) {
    if (!out) return set_error(MELA_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return set_error(MELA_ERR_DEVICE, "no HIP device available (libmela needs an MI355X)",
                         "mela:device");
    if (device < 0 || device >= ndev) return set_error(MELA_ERR_INVALID, "device index out of range");
    MELA_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    MELA_HIP(hipGetDeviceProperties(&prop, device));
    mela_ctx *ctx = new mela_ctx();
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    {
        const char *fp = std::getenv("MELA_FORCE_PAGEABLE");
        ctx->force_pageable = fp && *fp && std::strcmp(fp, "0") != 0;
    }
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    ctx->scratch.resize(SLOT_COUNT);
    // the allocation-free entry point's device array, once per device
    ctx->noalloc_scratch = noalloc_scratch_for_device(device, &ctx->noalloc_err);
    *out = ctx;
    return MELA_OK;
}

mela_status mela_ctx_default(mela_ctx **out) {
    static std::mutex mu;
    static mela_ctx *dflt = nullptr;
    if (!out) return set_error(MELA_ERR_INVALID, "null output pointer");
    std::lock_guard<std::mutex> lk(mu);
    if (!dflt) {
        mela_status st = mela_ctx_create(0, &dflt);
        if (st != MELA_OK) return st;
    }
    *out = dflt;
    return MELA_OK;
}

static void prof_clear(mela_ctx *ctx) {
    for (auto &pr : ctx->prof) {
        if (pr.first) (void)hipEventDestroy(pr.first);
        if (pr.second) (void)hipEventDestroy(pr.second);
    }
    ctx->prof.clear();
}

void mela_ctx_destroy(mela_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    prof_clear(ctx);
    for (auto &g : ctx->elim_graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    for (auto &s : ctx->scratch)
        if (s.ptr) (void)hipFree(s.ptr);
    for (auto &r : ctx->retired) (void)hipFree(r.first);
    pinned_release(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void *mela_ctx_stream(mela_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

mela_status mela_ctx_sync(mela_ctx *ctx) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    MELA_HIP(hipStreamSynchronize(ctx->stream));
    return MELA_OK;
}

mela_status mela_ctx_elim_counts(mela_ctx *ctx, unsigned long long *blocked,
                                 unsigned long long *fallback, unsigned long long *panel) {
    if (!ctx || !blocked || !fallback || !panel) return set_error(MELA_ERR_INVALID, "null argument");
    CtxLock lk(ctx->mu);
    *blocked = ctx->elim_blocked;
    *fallback = ctx->elim_fallback;
    *panel = ctx->elim_panel;
    return MELA_OK;
}

mela_status mela_ctx_profile_begin(mela_ctx *ctx) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    prof_clear(ctx);
    ctx->prof_on = true;
    return MELA_OK;
}

mela_status mela_ctx_profile_end(mela_ctx *ctx, double *total_ms, int *launches) {
    if (!ctx || !total_ms || !launches) return set_error(MELA_ERR_INVALID, "null argument");
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    ctx->prof_on = false;
    MELA_HIP(hipStreamSynchronize(ctx->stream));
    double sum = 0.0;
    int cnt = 0;
    for (auto &pr : ctx->prof) {
        if (!pr.first || !pr.second) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
            sum += ms;
            ++cnt;
        } else {
            (void)hipGetLastError();
        }
    }
    prof_clear(ctx);
    *total_ms = sum;
    *launches = cnt;
    return MELA_OK;
}

mela_status mela_dev_alloc(mela_ctx *ctx, size_t bytes, void **dptr) {
    if (!ctx || !dptr) return set_error(MELA_ERR_INVALID, "null argument");
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *dptr = nullptr;
        return set_error(MELA_ERR_NOMEM, "Out of memory.", "mela:memory");
    }
    return MELA_OK;
}

mela_status mela_dev_free(mela_ctx *ctx, void *dptr) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    if (dptr) {
        CtxLock lk(ctx->mu);
        MELA_HIP(hipSetDevice(ctx->device));
        MELA_HIP(hipStreamSynchronize(ctx->stream));
        MELA_HIP(hipFree(dptr));
    }
    return MELA_OK;
}

mela_status mela_copy_h2d(mela_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    if (bytes == 0) return MELA_OK;
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    MELA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    MELA_HIP(hipStreamSynchronize(ctx->stream));
    return MELA_OK;
}

mela_status mela_copy_d2h(mela_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    if (bytes == 0) return MELA_OK;
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    MELA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MELA_HIP(hipStreamSynchronize(ctx->stream));
    return MELA_OK;
}

mela_status mela_copy_d2d(mela_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    return launch_copy(ctx->stream, dst, src, bytes);
}

mela_status mela_bitmat_alloc(mela_ctx *ctx, size_t rows, size_t cols, mela_bitmat *out) {
    if (!ctx || !out) return set_error(MELA_ERR_INVALID, "null argument");
    CtxLock lk(ctx->mu);
    const size_t ld = words64(cols) ? words64(cols) : 1;
    void *p = nullptr;
    mela_status st = mela_dev_alloc(ctx, (rows ? rows : 1) * ld * 8, &p);
    if (st != MELA_OK) return st;
    st = launch_fill_zero(ctx->stream, p, (rows ? rows : 1) * ld * 8);
    if (st != MELA_OK) return st;
    out->data = (uint64_t *)p;
    out->rows = rows;
    out->cols = cols;
    out->ld = ld;
    return MELA_OK;
}

void mela_bitmat_free(mela_ctx *ctx, mela_bitmat *m) {
    if (!m || !m->data) return;
    (void)mela_dev_free(ctx, m->data);
    m->data = nullptr;
}

mela_status mela_gen_bernoulli(mela_ctx *ctx, uint64_t seed, mela_bitmat *M) {
    if (!ctx || !M || (!M->data && M->rows)) return set_error(MELA_ERR_INVALID, "null argument");
    if (M->ld < words64(M->cols)) return set_error(MELA_ERR_INVALID, "ld smaller than the row width");
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    return launch_gen_bernoulli(ctx->stream, seed, M->data, M->rows, M->cols, M->ld);
}

}  // extern "C"
