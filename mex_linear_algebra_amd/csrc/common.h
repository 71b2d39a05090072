// Internal declarations shared by the HIP sources of libmela.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "mela.h"

namespace mela {

// ------------------------------------------------------------- errors ----
// Thread-local last error (message + MATLAB id), returned by mela_last_error.
mela_status set_error(mela_status st, const std::string &msg, const char *id = "");
mela_status hip_fail(hipError_t e, const char *what);

#define MELA_HIP(call)                                             \
    do {                                                           \
        hipError_t e__ = (call);                                   \
        if (e__ != hipSuccess) return ::mela::hip_fail(e__, #call); \
    } while (0)

inline size_t words64(size_t n) { return (n + 63) / 64; }

// ------------------------------------------------------------ context ----
// Device scratch is taken from a per-context cache of blocks that grows to the
// largest request seen and is reused across calls (no hipMalloc on the
// steady-state path).
struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// A captured elimination sequence (all panel launches of one shape), replayed
// with hipGraphLaunch when the same shape and work buffers come back.
struct ElimGraph {
    size_t m = 0, n = 0, ld = 0;
    int variant = 0, rref = 0;
    // the work buffers the captured launches use: M, Q, meta, act, block state
    const void *bufs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int seen = 0;
    // panels (blocks, on the blocked path) the last uncaptured run needed
    // (the first one whose selection found the elimination complete, plus
    // one; all of them if none did)
    size_t panels_needed = 0;
    size_t g_end = 0;  // panels in the captured sequence
    // the blocked path (512-column blocks) missed a pivot on this shape once:
    // its calls take the panel path from then on
    bool blocked_failed = false;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
};

}  // namespace mela

struct mela_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // serialises calls on one context: every call that uses the context's
    // stream or scratch holds it (recursive: entry points call each other)
    std::recursive_mutex mu;
    std::vector<mela::Scratch> scratch;  // indexed by slot
    int cu_count = 256;
    hipEvent_t ev = nullptr;
    std::vector<mela::ElimGraph> elim_graphs;
    // graph capture of repeated eliminations; off for contexts that share the
    // process with other threads doing synchronous HIP calls (server workers)
    bool capture_ok = true;
    // a scratch slot that grows keeps its old block here instead of freeing
    // it: hipFree waits for every stream's work on the device, which a
    // server worker must not do in the middle of a request (freed by
    // ctx_release_scratch and at destruction)
    bool defer_free = false;
    std::vector<std::pair<void *, size_t>> retired;
    // event pairs around the dominant kernel of each product while profiling
    bool prof_on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof;
    // pinned host staging of the host-array entry points (grow-only)
    void *pin = nullptr;
    size_t pin_bytes = 0;
    bool pin_is_malloc = false;  // hipHostMalloc failed: ordinary host memory
    // staging in ordinary host memory from the start (MELA_FORCE_PAGEABLE=1 at
    // context creation): the path taken when no pinned memory can be had
    bool force_pageable = false;
    // eliminations by path (mela_ctx_elim_counts)
    unsigned long long elim_blocked = 0, elim_fallback = 0, elim_panel = 0;
    // this device's allocation-free scratch (mela_rank_packed_noalloc), or
    // nullptr with the reason in noalloc_err
    void *noalloc_scratch = nullptr;
    std::string noalloc_err;
};

namespace mela {

// The context lock (mela_ctx::mu), held for the whole of a call.
using CtxLock = std::lock_guard<std::recursive_mutex>;

// Records an event on the context stream before (start) or after the
// dominant kernel of a product when profiling is on; never used inside a
// captured graph.
inline void prof_mark(mela_ctx *ctx, bool start) {
    if (!ctx->prof_on) return;
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    (void)hipEventRecord(e, ctx->stream);
    if (start) {
        ctx->prof.push_back({e, nullptr});
    } else if (!ctx->prof.empty() && ctx->prof.back().second == nullptr) {
        ctx->prof.back().second = e;
    } else {
        (void)hipEventDestroy(e);
    }
}

// Returns a device buffer of at least `bytes` for `slot` (kept across calls).
mela_status scratch_get(mela_ctx *ctx, int slot, size_t bytes, void **out);

// Device bytes the context's scratch cache holds, and its release (waits for
// the context stream; drops captured eliminations that point into it).
size_t ctx_scratch_bytes(mela_ctx *ctx);
void ctx_release_scratch(mela_ctx *ctx);

// Returns the context's host staging buffer (pinned, or ordinary host memory
// when no pinned memory can be had), at least `bytes` long, once the context
// stream is idle (no earlier copy still reads or writes it).
mela_status pinned_get(mela_ctx *ctx, size_t bytes, uint8_t **out);
void pinned_release(mela_ctx *ctx);

enum ScratchSlot {
    SLOT_A = 0, SLOT_B, SLOT_C, SLOT_BT, SLOT_IN_A, SLOT_IN_B, SLOT_OUT,
    SLOT_ELIM_M, SLOT_ELIM_Q, SLOT_ELIM_META, SLOT_ELIM_ACT, SLOT_ELIM_TAB,
    SLOT_NULL_Z, SLOT_XA, SLOT_XB,
    SLOT_BATCH_DESC, SLOT_BATCH_WORK, SLOT_BATCH_Z, SLOT_BATCH_RANK, SLOT_NULL_TAB, SLOT_DENSITY,
    SLOT_ELIM_BLK,
    SLOT_COUNT
};

// Device-side choice between the row-XOR product and a dense one (AUTO on
// large products): a popcount kernel sums the set bits of A into *ones; the
// row-XOR kernel runs only if *ones < thr, the dense path's kernels only if
// not.  ones == nullptr: no gate, the kernel always runs.
// The count is spread over GATE_SLOTS words (workgroup b adds to slot
// b % GATE_SLOTS: one address for all of them serialised the atomics).
constexpr int GATE_SLOTS = 16;
struct Gate {
    const unsigned long long *ones = nullptr;  // GATE_SLOTS partial counts
    unsigned long long thr = 0;
};
__device__ __forceinline__ bool gate_skip(const Gate &g, bool sparse_path) {
    if (!g.ones) return false;
    unsigned long long t = 0;
#pragma unroll
    for (int i = 0; i < GATE_SLOTS; ++i) t += g.ones[i];
    return (t < g.thr) != sparse_path;
}

// Device-resident serving loop (serve.cpp, kernel in batch.hip).  The slots,
// the done words and the tail live in pinned host memory, which the kernel
// reads and writes with plain system-scope loads and stores (no PCIe
// read-modify-write); the take counter `head` lives in device memory.  The
// host fills a slot, then publishes it by raising tail; a workgroup takes
// slot `head` by compare-and-swap only while head < tail, eliminates the
// matrix, and stores seq + 1 into the slot's done word once its rank and
// null basis are visible system-wide.  Tail's top bit asks every workgroup
// to end.
struct ServeSlot {
    const uint64_t *A;  // device input, m x lda words
    uint64_t *work;     // device work: m x wc working rows, then min(m,n) x wc pivot rows
    uint64_t *Z;        // device null basis (n x wc words), null for rank requests
    uint32_t m, n, lda, op;  // op: MELA_OP_RANK or MELA_OP_NULL
    uint32_t mode, pad;      // 0 GE, 1 M4RI sweep
};
struct ServeDone {
    uint32_t rank;
    uint32_t pad;
    uint64_t seq1;  // sequence number + 1 once finished
};
constexpr uint64_t SERVE_STOP = 1ull << 63;  // in tail: end the loop
// Workgroups of one serving-loop launch and the launcher: idle_ticks of the
// 100 MHz wall clock without a published slot, or a launch life_ticks old,
// end a workgroup between requests.
int serve_loop_grid(mela_ctx *ctx);
mela_status launch_serve_loop(hipStream_t s, int grid, ServeSlot *slots, ServeDone *done,
                              const uint64_t *tail, uint64_t *head, uint32_t cap, uint64_t idle_ticks,
                              uint64_t life_ticks);

// ------------------------------------------------- kernel launchers ----
// pack.hip
mela_status launch_pack(hipStream_t s, const void *src, int cls, size_t rows,
                        size_t cols, uint64_t *dst, size_t ld);
mela_status launch_unpack(hipStream_t s, const uint64_t *src, size_t ld,
                          size_t rows, size_t cols, uint8_t *dst);
mela_status launch_transpose(hipStream_t s, const uint64_t *src, size_t lds,
                             size_t rows, size_t cols, uint64_t *dst, size_t ldd);
mela_status launch_gen_bernoulli(hipStream_t s, uint64_t seed, uint64_t *dst,
                                 size_t rows, size_t cols, size_t ld);
mela_status launch_fill_zero(hipStream_t s, void *dst, size_t bytes);
mela_status launch_pack_csc(hipStream_t s, const uint64_t *jc, const uint64_t *ir,
                            const void *vals, int vcls, size_t nnz, size_t m, size_t n,
                            uint64_t *dst, size_t ld, unsigned *err);
mela_status launch_copy(hipStream_t s, void *dst, const void *src, size_t bytes);

// matmul.hip
mela_status matmul_device(mela_ctx *ctx, const uint64_t *A, size_t lda,
                          const uint64_t *B, size_t ldb, uint64_t *C, size_t ldc,
                          size_t m, size_t k, size_t n, int algo);

// mfma.hip
// gate / count / sparse_path: the AUTO product's density choice.  With
// count set, the call adds the set bits of A to *count (zeroed by the
// caller) on its way and calls sparse_path(gate) once the count is complete,
// to enqueue the row-XOR product; the dense kernels run behind the gate.
mela_status matmul_mfma(mela_ctx *ctx, const uint64_t *A, size_t lda, const uint64_t *B,
                        size_t ldb, uint64_t *C, size_t ldc, size_t m, size_t k, size_t n,
                        bool fp4, Gate gate = Gate(), unsigned long long *count = nullptr,
                        const std::function<mela_status(Gate)> &sparse_path = nullptr);
// the GATE_SLOTS counts at ones += set bits of the rows x kw words of A (enqueued)
mela_status launch_popcount(mela_ctx *ctx, const uint64_t *A, size_t lda, size_t rows, size_t kw,
                            unsigned long long *ones);

// elim.hip
struct ElimResult {
    size_t rank = 0;
    std::vector<int> pivots;  // filled when want_pivots
};
// Caller-provided work buffers of one elimination (the layout of
// elim_layout); with them eliminate() takes nothing from the context's
// scratch and neither captures nor replays a graph.
struct ElimBufs {
    uint64_t *M = nullptr;   // working copy, m x ld words
    uint64_t *Q = nullptr;   // pivot rows, min(m, n) x ld words
    void *meta = nullptr;    // metadata, pivot columns, sweep tags, panel words
    uint32_t *act = nullptr; // active-row list, m entries
};
// Byte offsets of the four buffers in one region, and its size.
struct ElimLayout {
    size_t ld = 0;  // words per working row (a multiple of 16)
    size_t meta_bytes = 0, act_off = 0, m_off = 0, q_off = 0, total = 0;
};
ElimLayout elim_layout(size_t m, size_t n);
// Runs elimination on a packed copy owned by the call.  mode_rref: full RREF
// (also reduces earlier pivot rows); Q receives pivot rows in order.
// fixed: work buffers given by the caller (see ElimBufs), else the context's.
mela_status eliminate(mela_ctx *ctx, const uint64_t *A, size_t lda, size_t m,
                      size_t n, int variant, bool rref, bool want_pivots,
                      ElimResult *res, uint64_t **Q_out, size_t *ldq_out,
                      const ElimBufs *fixed = nullptr);
// Address of this device's allocation-free scratch array (elim.hip), looked
// up once per device; the device must be current.  nullptr and a message on
// failure.
void *noalloc_scratch_for_device(int device, std::string *err);
// packed null basis (n x f, row-major, ldz words, device) from the RREF rows;
// returns with the stream idle
mela_status launch_null_packed(mela_ctx *ctx, const uint64_t *Q, size_t ldq, const int *rowmap,
                               const std::vector<int> &piv, size_t n, size_t f, uint64_t *Z,
                               size_t ldz);

}  // namespace mela
