/*
 * mela.h -- C ABI of the MI355X (gfx950) GF(2) linear-algebra library.
 *
 * Drop-in for the six MEX functions of HCPhy/mex_linear_algebra.  Each of the
 * reference's gateways (void mexFunction(int nlhs, mxArray *plhs[], int nrhs,
 * const mxArray *prhs[])) validates its mxArray arguments, bit-packs them,
 * runs a CPU core and materialises a logical/double result.  Here the
 * validation order and message text are kept, the packing, elimination and
 * output materialisation run as HIP kernels, and errors come back as status
 * codes plus a thread-local message instead of a longjmp out of the call.
 *
 * Matrices handed to the six entry points are MATLAB column-major arrays:
 * element (i,j) of an m x n array sits at index i + j*m.  A "logical" array
 * is one byte per element (nonzero = true); a "double" array is read as
 * ((int)x) & 1 (x86-64 semantics for NaN/Inf/out-of-range: parity 0; the
 * reference leaves these undefined, and its arm64 builds saturate instead, so
 * +Inf and x >= 2^31 would have parity 1 there).  Complex doubles are passed
 * by their real parts (the MEX gateways do this under either MATLAB array
 * API, as the reference's mxGetData / mxGetPr reads do).
 *
 * All functions are safe to call from several host threads on distinct
 * contexts; one context serialises its own calls.  Passing ctx == NULL to an
 * entry point uses the library's default context (device 0, created on first
 * use, after argument validation has passed).
 */
#ifndef MELA_H
#define MELA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mela_status {
    MELA_OK = 0,
    MELA_ERR_SPARSE = 1,    /* sparse input (reference rejects, matmul_gf2.c:78-79) */
    MELA_ERR_DIM = 2,       /* inner dimensions differ (matmul_gf2.c:84) */
    MELA_ERR_TYPE = 3,      /* neither logical nor double (matmul_gf2.c:86-88) */
    MELA_ERR_NOMEM = 4,     /* host or device allocation failed */
    MELA_ERR_INVALID = 5,   /* null pointer / size overflow / bad descriptor */
    MELA_ERR_DEVICE = 6,    /* HIP runtime error (message names the call) */
    MELA_ERR_ADMISSION = 7, /* request larger than the server's HBM budget */
    MELA_ERR_CANCELLED = 8, /* request cancelled or server shut down */
    MELA_ERR_TIMEOUT = 9,   /* wait deadline passed */
    MELA_ERR_ARGS = 10      /* wrong number of inputs/outputs (MEX gateways) */
} mela_status;

/* Input class: the two mxClassIDs the reference accepts, plus OTHER for any
 * other class.  OR with MELA_SPARSE for a sparse array. */
enum {
    MELA_LOGICAL = 0,
    MELA_DOUBLE = 1,
    MELA_OTHER = 2,
    MELA_SPARSE = 0x100
};

typedef struct mela_ctx mela_ctx;

/* ----------------------------------------------------------- context ---- */
mela_status mela_ctx_create(int device, mela_ctx **ctx);
/* the shared default context (device 0), created on first call */
mela_status mela_ctx_default(mela_ctx **ctx);
void mela_ctx_destroy(mela_ctx *ctx);
/* HIP stream (hipStream_t) every kernel of this context runs on */
void *mela_ctx_stream(mela_ctx *ctx);
mela_status mela_ctx_sync(mela_ctx *ctx);
/* Kernel timing (no reference counterpart; for benchmarks and profiling):
   between begin and end, every mela_matmul_packed call on this context
   records HIP events around its dominant kernel (k_andpop, k_m4rm or the
   MFMA product kernel); end synchronises the stream and returns the summed
   kernel time and the number of timed launches. */
mela_status mela_ctx_profile_begin(mela_ctx *ctx);
mela_status mela_ctx_profile_end(mela_ctx *ctx, double *total_ms, int *launches);
/* Eliminations run by this context, by path (no reference counterpart; for
   tests and benchmarks): blocked = finished on the blocked path, fallback =
   blocked attempts redone on the panel path, panel = panel path only. */
mela_status mela_ctx_elim_counts(mela_ctx *ctx, unsigned long long *blocked,
                                 unsigned long long *fallback, unsigned long long *panel);
/* message and MATLAB identifier of the last failure on this thread */
const char *mela_last_error(void);
const char *mela_last_error_id(void);
const char *mela_version(void);

/* ------------------------------------------- the six reference entry points
 *
 * mela_matmul_gf2 / mela_matmul_m4ri replace mexFunction at
 * src/mela_matmul_gf2.c:72 and src/mela_matmul_m4ri.c:140:
 *   C (m x n logical, caller-allocated m*n bytes) = A (m x k) * B (k2 x n)
 * Checks in reference order: sparse ("Sparse matrices not supported. Convert
 * to full using full()."), k == k2 ("Inner dimensions must agree."), classes
 * ("Inputs must be logical or double matrices.").  Both run MELA_ALGO_AUTO
 * (the product is the same whatever the algorithm; MELA_MATMUL_ALGO=andpop,
 * m4rm, mfma_i8 or mfma_fp4 in the environment forces one). */
mela_status mela_matmul_gf2(mela_ctx *ctx, const void *A, int a_class,
                            size_t m, size_t k, const void *B, int b_class,
                            size_t k2, size_t n, uint8_t *C);
mela_status mela_matmul_m4ri(mela_ctx *ctx, const void *A, int a_class,
                             size_t m, size_t k, const void *B, int b_class,
                             size_t k2, size_t n, uint8_t *C);

/* mela_rank_gf2 replaces src/mela_rank_gf2.c:126 (class message "Input must
 * be a logical or double matrix.", id bitrank_mex:inputNotLogical);
 * mela_rank_m4ri replaces src/mela_rank_m4ri.c:286 ("Input must be double or
 * logical.").  The rank is returned as a double, like mxCreateDoubleScalar.
 * Sparse input is rejected with the matmul message (the reference gateways do
 * not check and would read a sparse array's storage as dense data). */
mela_status mela_rank_gf2(mela_ctx *ctx, const void *A, int a_class,
                          size_t m, size_t n, double *rank);
mela_status mela_rank_m4ri(mela_ctx *ctx, const void *A, int a_class,
                           size_t m, size_t n, double *rank);

/* mela_null_gf2 / mela_null_m4ri replace src/mela_null_gf2.c:142 and
 * src/mela_null_m4ri.c:251.  The n x (n - rank) logical basis Z is the
 * canonical one (Z(free,:) = I, A*Z = 0, free columns ascending); its size is
 * known only after elimination, so the library calls alloc(user, n, nullity)
 * once and fills the column-major bytes it returns (the MEX gateway passes a
 * function that creates the mxArray, as null_gf2.c:208 does).  Messages:
 * sparse (as matmul), "H must be double or logical.", "Out of memory.". */
typedef uint8_t *(*mela_alloc_logical_fn)(void *user, size_t rows, size_t cols);
mela_status mela_null_gf2(mela_ctx *ctx, const void *H, int h_class,
                          size_t m, size_t n, mela_alloc_logical_fn alloc,
                          void *user, size_t *nullity);
mela_status mela_null_m4ri(mela_ctx *ctx, const void *H, int h_class,
                           size_t m, size_t n, mela_alloc_logical_fn alloc,
                           void *user, size_t *nullity);

/* Reduced row echelon form with pivot columns, the semantics of the MATLAB
 * comparator the reference ships, gf2null/gf2rref.m:49-73 (no MEX
 * counterpart): R (m x n logical, caller-allocated m*n bytes, column-major)
 * receives the RREF with its rank nonzero rows on top; jb (NULL or min(m,n)
 * doubles) receives the 1-based pivot columns in ascending order; *rank =
 * length(jb).  Messages: sparse (as matmul), "Input must be a logical or
 * double matrix.". */
mela_status mela_rref_gf2(mela_ctx *ctx, const void *A, int a_class, size_t m, size_t n,
                          uint8_t *R, double *jb, size_t *rank);

/* ------------------------------------------- device-resident packed forms
 *
 * A packed matrix lives in device memory, row-major: row i is ld 64-bit
 * words, bit j of row i is bit (j & 63) of word (j >> 6) (the reference's
 * packed layout, matmul_gf2.c:33-36).  Bits past column cols must be zero. */
typedef struct mela_bitmat {
    uint64_t *data; /* device pointer */
    size_t rows, cols, ld;
} mela_bitmat;

/* Product algorithms.  AUTO: fp4 MFMA from 2^33 bit-products on, AND+parity
 * below; for products that would take over ~40 us it counts the set bits of
 * A on the device and takes ROWXOR when A is sparse enough (one B-row read
 * per set bit costs less than the dense kernel). */
enum {
    MELA_ALGO_AUTO = 0,
    MELA_ALGO_ANDPOP = 1, /* AND + parity over a transposed B (v_bitop3) */
    MELA_ALGO_M4RM = 2,   /* Method of Four Russians, Gray-code tables in LDS */
    MELA_ALGO_MFMA_I8 = 3, /* exact int8 MFMA, i32 accumulate, & 1 */
    MELA_ALGO_MFMA_FP4 = 4, /* exact fp4 (e2m1) MFMA, f32 accumulate, & 1; k < 2^24.
                               Grids of >= one 256 x 256 tile per CU with k < 2^23 read the
                               packed bits as fp4 planes in registers; others (or any, with
                               MELA_MFMA_EXPANDED=1 in the environment) expand the operands
                               to e2m1 nibbles first */
    MELA_ALGO_ROWXOR = 5  /* C_i ^= B_k for each set bit A(i,k) (matmul_gf2.c:127-186) */
};

mela_status mela_dev_alloc(mela_ctx *ctx, size_t bytes, void **dptr);
mela_status mela_dev_free(mela_ctx *ctx, void *dptr);
mela_status mela_copy_h2d(mela_ctx *ctx, void *dst, const void *src, size_t bytes);
mela_status mela_copy_d2h(mela_ctx *ctx, void *dst, const void *src, size_t bytes);
/* device-to-device copy by a plain kernel (16-byte aligned; enqueued only) */
mela_status mela_copy_d2d(mela_ctx *ctx, void *dst, const void *src, size_t bytes);
mela_status mela_bitmat_alloc(mela_ctx *ctx, size_t rows, size_t cols, mela_bitmat *out);
void mela_bitmat_free(mela_ctx *ctx, mela_bitmat *m);

/* Bernoulli(1/2) bits: word (i,w) = splitmix64(splitmix64(k1 + i*ceil(cols/64) + w) ^ k2),
 * k1 = splitmix64(seed), k2 = splitmix64(seed ^ 0xD1B54A32D192ED03) */
mela_status mela_gen_bernoulli(mela_ctx *ctx, uint64_t seed, mela_bitmat *M);

/* pack a column-major logical/double host-or-device array into M (device) */
mela_status mela_pack_dev(mela_ctx *ctx, const void *src_dev, int src_class,
                          mela_bitmat *M);
/* unpack M into a column-major logical device array of rows*cols bytes */
mela_status mela_unpack_dev(mela_ctx *ctx, const mela_bitmat *M, uint8_t *dst_dev);

/* C = A * B over GF(2) (C->rows == A->rows, C->cols == B->cols,
 * A->cols == B->rows).  Enqueued on the context stream; returns at once. */
mela_status mela_matmul_packed(mela_ctx *ctx, const mela_bitmat *A,
                               const mela_bitmat *B, mela_bitmat *C, int algo);

/* Elimination variants.  Outputs are identical -- rank, the reduced row
 * echelon form and the canonical null basis are unique -- so the six
 * host-array entry points use AUTO whichever reference algorithm they stand
 * for.  The panel path eliminates 64 columns per step and sweeps every
 * remaining row with the panel's pivot rows: GE applies them one by one, M4RI
 * through 4-bit tables, PANEL decides per row batch (the walk for
 * multipliers with few set bits, the tables otherwise).  The blocked path
 * factors 512 columns at a time from the head of the remaining rows and
 * applies them to every other row in one matrix-core product; when a block's
 * pivots are not all in its head (sparse or structured matrices) the call
 * redoes the elimination on the panel path, and the context keeps that
 * shape on the panel path.  AUTO: blocked from 2048 x 2048 on, else PANEL;
 * BLOCKED: the blocked path at any size. */
enum {
    MELA_ELIM_GE = 0,
    MELA_ELIM_M4RI = 1,
    MELA_ELIM_AUTO = 2,
    MELA_ELIM_BLOCKED = 3,
    MELA_ELIM_PANEL = 4
};

/* rank of A (A is left unchanged; work memory comes from the context) */
mela_status mela_rank_packed(mela_ctx *ctx, const mela_bitmat *A, int variant,
                             size_t *rank);
/* Allocation-free rank.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same result as mela_rank_packed,
 * but the work memory (working copy, pivot rows, metadata, active list) is a
 * fixed 32 MiB device array of the library, one per GPU, held until the
 * process ends; the call makes no hipMalloc / hipFree and no other device
 * allocation.  A matrix whose work memory exceeds 32 MiB (about
 * 8 * ceil16(ceil(cols/64)) * (rows + min(rows, cols)) bytes plus
 * 4 * rows + 25 KiB: 16384 x 8192 fits, 16384 x 16384 does not) is
 * rejected with MELA_ERR_INVALID.  Calls on one GPU share that array: the library runs
 * them one at a time within a process (they synchronise their stream), and
 * the array keeps the last call's contents. */
mela_status mela_rank_packed_noalloc(mela_ctx *ctx, const mela_bitmat *A, int variant,
                                     size_t *rank);
/* reduced row echelon form: R (min(m,n) x cols, device, may be NULL) gets the
 * rank nonzero rows in pivot order; pivots (host, min(m,n) ints, may be NULL)
 * the pivot columns ascending. */
mela_status mela_rref_packed(mela_ctx *ctx, const mela_bitmat *A, int variant,
                             mela_bitmat *R, int *pivots, size_t *rank);
/* canonical null basis, packed: Z is cols x nullity, allocated by the call
 * (free with mela_bitmat_free). */
mela_status mela_null_packed(mela_ctx *ctx, const mela_bitmat *A, int variant,
                             mela_bitmat *Z, size_t *nullity);

/* Sparse input (no reference counterpart: the reference rejects sparse
 * arrays, matmul_gf2.c:78-79, null_gf2.c:151-153).  Packs an m x n matrix in
 * MATLAB's compressed sparse column form into A (allocated by the caller,
 * rows = m, cols = n; its contents are replaced): column j's stored entries
 * are rows ir[jc[j] .. jc[j+1]-1] (0-based, uint64 like mwIndex).  values:
 * NULL (every stored entry is 1) or nnz logical bytes / doubles
 * (values_class MELA_LOGICAL / MELA_DOUBLE; the dense parity rule).  jc must
 * start at 0 and not decrease; row indices >= m are an error. */
mela_status mela_pack_csc(mela_ctx *ctx, const uint64_t *jc, const uint64_t *ir,
                          const void *values, int values_class, size_t m, size_t n,
                          mela_bitmat *A);

/* ------------------------------------------------ batched rank / null space
 *
 * No reference counterpart (the reference handles one matrix per MEX call):
 * rank, RREF and canonical null bases of up to millions of uneven matrices
 * (each at most 4096 x 8192) in one persistent launch.  mela_batch_create
 * plans the batch (work arenas, largest-first queue); mela_batch_run enqueues
 * it (one hipGraph launch with MELA_BATCH_GRAPH); mela_batch_results waits and
 * returns ranks and null-basis views (device memory owned by the batch, valid
 * until mela_batch_destroy or the next run). */
typedef struct mela_batch mela_batch;
enum {
    MELA_BATCH_NULL = 1,  /* also compute the packed null bases */
    MELA_BATCH_GE = 2,    /* start every matrix in GE mode (default M4RI) */
    MELA_BATCH_GRAPH = 4, /* capture the launch in a hipGraph and replay it */
    MELA_BATCH_CTX_SCRATCH = 8 /* borrow the context's cached arenas (only one
                                  such batch may exist per context at a time) */
};
mela_status mela_batch_create(mela_ctx *ctx, const mela_bitmat *mats, size_t count,
                              int flags, mela_batch **batch);
/* per-matrix elimination mode: 0 = GE (XOR each pivot row), 1 = M4RI with
 * k = 4 (16 lookups per row word), 2 = M4RI with k = 8 (8 lookups per row
 * word; takes effect with 8-word tiles, runs as k = 4 with wider tiles).
 * The reference's M4RI k is a compile-time constant (mela_rank_m4ri.c:18,
 * mela_null_m4ri.c:28); here it is chosen per matrix. */
mela_status mela_batch_set_modes(mela_batch *batch, const int *modes);
/* update tile width in 64-bit words: 8 (the only width with k = 8 tables),
 * 16 (default) or 32 */
mela_status mela_batch_set_tile(mela_batch *batch, int tile_words);
/* the plan the next run uses: tile width, per-matrix modes (as set: a 2
 * reported with 16- or 32-word tiles runs as 1) and workgroups per matrix
 * (each output may be NULL; modes and groups hold count entries) */
mela_status mela_batch_get_plan(mela_batch *batch, int *tile_words, int *modes, int *groups);
/* per-matrix workgroups (1..8) that eliminate the matrix together: the
 * tiles of each panel are dealt out among them, one device-scope barrier per
 * panel.  Every workgroup of the launch is resident at once, so the group
 * barriers only wait for running partners; on a device shared with other
 * work a barrier that waits over 0.5 s fails the run (MELA_ERR_DEVICE from
 * mela_batch_results).  Default 1. */
mela_status mela_batch_set_groups(mela_batch *batch, const int *groups);
/* times candidate plans on the batch itself and keeps the fastest: tile
 * width x per-matrix mode (GE, or M4RI with k = 4 or 8), then workgroups
 * per matrix for the largest matrices; *best_ms receives its time (may be
 * NULL); mela_batch_get_plan reports the choice */
mela_status mela_batch_autotune(mela_batch *batch, float *best_ms);
mela_status mela_batch_run(mela_batch *batch);
mela_status mela_batch_results(mela_batch *batch, uint32_t *ranks, mela_bitmat *nulls);
void mela_batch_destroy(mela_batch *batch);

/* ---------------------------------------------------- batched products
 *
 * No reference counterpart (the reference multiplies one pair per MEX call,
 * matmul_gf2.c:72): C_i = A_i * B_i for thousands of uneven packed device
 * products in one persistent launch.  mela_matmul_batch_create validates the
 * shapes (A_i.cols == B_i.rows, C_i is A_i.rows x B_i.cols), counts the set
 * bits of every A_i and gives each item the fp4 matrix cores (128 x 128
 * tiles; AND+parity when k >= 2^23) or the row-XOR product (sparse A_i),
 * whichever the measured kernel rates say is cheaper;
 * mela_matmul_batch_run enqueues all products (one hipGraph launch with
 * MELA_BATCH_GRAPH) on the context stream.  The C_i receive their products
 * (words past each row's last column are left as they are); the inputs must
 * stay valid while a run is in flight. */
typedef struct mela_matmul_batch mela_matmul_batch;
mela_status mela_matmul_batch_create(mela_ctx *ctx, const mela_bitmat *A, const mela_bitmat *B,
                                     const mela_bitmat *C, size_t count, int flags,
                                     mela_matmul_batch **batch);
/* per-item engine: MELA_ALGO_ANDPOP, MELA_ALGO_MFMA_FP4 (k < 2^23) or
 * MELA_ALGO_ROWXOR (set, or read the planned choice) */
mela_status mela_matmul_batch_set_algos(mela_matmul_batch *batch, const int *algos);
mela_status mela_matmul_batch_get_algos(mela_matmul_batch *batch, int *algos);
mela_status mela_matmul_batch_run(mela_matmul_batch *batch);
void mela_matmul_batch_destroy(mela_matmul_batch *batch);

/* ---------------------------------------------------------------- serving
 *
 * No reference counterpart.  A server admits mixed rank / null-space / matmul
 * requests on device-resident packed inputs in arrival order while the
 * device memory they need fits hbm_budget_bytes (everything the server holds
 * counts: its buffer pool, its workers' scratch caches, outputs not yet
 * released, and an upper bound for each running job).  Rank and null
 * requests that one workgroup finishes quickly go to a device-resident loop,
 * a persistent kernel that takes each request as soon as it is published
 * (at most max_batch of them on the device at once, up to 4096); larger ones
 * and all products run on `workers` streams with the single-matrix entry
 * points.  Latency percentiles are kept.  Inputs must stay valid until the
 * request completes; outputs (null basis or product) stay on the device
 * until mela_server_release. */
enum { MELA_OP_RANK = 0, MELA_OP_NULL = 1, MELA_OP_MATMUL = 2 };
typedef struct mela_request {
    int op;          /* MELA_OP_* */
    int variant;     /* MELA_ELIM_* for rank/null, MELA_ALGO_* for matmul */
    mela_bitmat A;   /* device input */
    mela_bitmat B;   /* second factor (matmul only) */
} mela_request;
typedef struct mela_result {
    mela_status status;
    size_t rank;       /* rank (rank and null requests) */
    mela_bitmat out;   /* null basis (cols x nullity) or product; device */
    double latency_ms; /* submit -> completion */
} mela_result;
typedef struct mela_server_stats {
    uint64_t submitted, completed, rejected, cancelled, failed;
    double p50_ms, p99_ms, mean_ms;
    size_t hbm_in_use, hbm_peak;
    uint64_t jobs;              /* requests run by the workers */
    uint64_t batched_requests;  /* requests run by the device-resident loop */
    uint64_t loop_launches;     /* starts of the loop kernel (an instance lives <= 1 ms) */
} mela_server_stats;
typedef struct mela_server mela_server;
mela_status mela_server_create(int device, size_t hbm_budget_bytes, int workers,
                               int max_batch, mela_server **server);
/* queues a request; MELA_ERR_ADMISSION if it alone exceeds the budget */
mela_status mela_server_submit(mela_server *server, const mela_request *req,
                               uint64_t *ticket);
/* waits (timeout_s < 0: forever); MELA_ERR_TIMEOUT / MELA_ERR_CANCELLED */
mela_status mela_server_wait(mela_server *server, uint64_t ticket, double timeout_s,
                             mela_result *result);
mela_status mela_server_cancel(mela_server *server, uint64_t ticket);
/* frees a finished request's device output and forgets the ticket */
mela_status mela_server_release(mela_server *server, uint64_t ticket);
mela_status mela_server_stats_get(mela_server *server, mela_server_stats *stats);
mela_status mela_server_reset_stats(mela_server *server);
void mela_server_destroy(mela_server *server);

#ifdef __cplusplus
}
#endif
#endif /* MELA_H */
