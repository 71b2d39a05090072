// Batched uneven GF(2) products C_i = A_i * B_i in one persistent launch (no
// reference counterpart: the reference multiplies one pair per MEX call,
// matmul_gf2.c:72 / matmul_m4ri.c:140).
//
// Planning (mela_matmul_batch_create) counts the set bits of every A_i on the
// device and gives each item the cheaper of two engines by the measured
// rates of the kernels: the fp4 matrix cores over A_i and B_i^T (cost m n k
// bit operations at their rate; AND+parity over B_i^T when k >= 2^23) or the
// row-XOR product (one B_i-row read per set bit of A_i, the reference's own
// GE product, matmul_gf2.c:127-186).  Items are queued largest first;
// each item is cut into output tiles (64 x 256 bits for AND+parity, 128 x
// 128 for the matrix cores, 4 rows x 64 words for row XOR).  A run is three
// launches, captured once in a hipGraph: reset of the queue counter, one
// batched B^T transpose for the AND+parity and matrix-core items, and the
// persistent product kernel whose workgroups take tiles from the queue until
// it is empty.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "mfma_device.h"
#include "product_device.h"
#include "transpose_device.h"

namespace mela {

struct MMItem {
    const uint64_t *A;
    const uint64_t *B;
    uint64_t *C;
    uint64_t bt_off;      // words: B^T in the arena (AND+parity items)
    uint64_t tile_begin;  // first queue tile of the item (prefix in queue order)
    uint64_t tb_begin;    // first transpose block, 64 k-rows x MT_WORDS column words (prefix in queue order)
    uint32_t m, k, n;
    uint32_t lda, ldb, ldc, ldbt;
    uint32_t algo;        // MM_ANDPOP, MM_ROWXOR or MM_MFMA
    uint32_t tiles_x;     // tiles along the rows of C
    uint32_t pad;
};

namespace {

enum { MM_ANDPOP = 0, MM_ROWXOR = 1, MM_MFMA = 2 };
constexpr int MF_TILE = 128;  // rows and columns of C per matrix-core tile
constexpr uint32_t MF_KMAX = 1u << 23;  // the tile's accumulators start at 2^23
constexpr int MT_WORDS = 4;             // column words per wave of the B^T transpose (8: no faster)

using dev::AP_COLS;
using dev::AP_ROWS;
using dev::RX_WAVES;

// Largest index j with key(j) <= t over items[0, n) (key non-decreasing,
// key(0) = 0 <= t): the item whose range [key(j), key(j+1)) holds t, items
// with empty ranges skipped.
template <typename Key>
__device__ __forceinline__ int find_item(int n, uint64_t t, Key key) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (key(mid) <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_mm_reset(unsigned long long *counter) {
    if (threadIdx.x == 0) *counter = 0ull;
}

// B_i^T (n x ldbt words) of every AND+parity and matrix-core item: one wave
// per 64 k-rows x MT_WORDS column words (lane = k row, one 64 x 64 bit
// transpose per word; one column word per wave ran 3x longer on its strided
// 8-byte loads),
// grid-strided over all items' blocks.
__global__ void __launch_bounds__(256)
k_mm_transpose(const MMItem *__restrict__ items, int nitems, uint64_t total_blocks,
               uint64_t *__restrict__ arena) {
    const int lane = threadIdx.x & 63;
    for (uint64_t b = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); b < total_blocks;
         b += (uint64_t)gridDim.x * 4) {
        const int ii = find_item(nitems, b, [&](int j) { return items[j].tb_begin; });
        const MMItem it = items[ii];
        const uint64_t lb = b - it.tb_begin;
        const uint32_t nbk = (it.k + 63) / 64;
        const uint32_t bk = (uint32_t)(lb % nbk), bn0 = MT_WORDS * (uint32_t)(lb / nbk);
        const uint32_t nw = (it.n + 63) / 64;
        const uint32_t i = bk * 64 + lane;
        uint64_t v[MT_WORDS];
#pragma unroll
        for (int q = 0; q < MT_WORDS; ++q) {
            const uint32_t bn = bn0 + q;
            v[q] = i < it.k && bn < nw ? it.B[(size_t)i * it.ldb + bn] : 0ull;
            if (bn == nw - 1 && (it.n & 63)) v[q] &= (1ull << (it.n & 63)) - 1ull;
        }
#pragma unroll
        for (int q = 0; q < MT_WORDS; ++q) {
            const uint64_t t = dev::transpose64(v[q], lane);
            const uint32_t c = (bn0 + q) * 64 + lane;
            if (c < it.n) arena[it.bt_off + (size_t)c * it.ldbt + bk] = t;
        }
    }
}

// One 128 x 128 tile of C_i on the fp4 matrix cores: four waves of 64 x 64
// (2 x 2 MFMA tiles), fragments read straight from the row-major packed rows
// of A_i and of B_i^T (the arena) with the stage-tiled layout's map (stage
// s, k half h of a lane: words 4s + 2h and 4s + 2h + 1 of its row), so A and
// B^T agree on which k each fp4 element holds; the next stage's fragments
// load before the current stage's MFMAs.  Accumulators start at 2^23 (the
// planner keeps k < 2^23) and the epilogue packs parities as the direct
// product kernel does.
__device__ __forceinline__ void mfma_tile(const MMItem &it, const uint64_t *__restrict__ bt, int row0,
                                          int col0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int half = lane >> 5, l32 = lane & 31;
    const int m = (int)it.m, n = (int)it.n;
    const int kw = (int)((it.k + 63) / 64), nst = (int)((it.k + 255) / 256);
    const uint64_t *ar[2], *br[2];
    bool av[2], bv[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int r = row0 + wr * 64 + 32 * t + l32, c = col0 + wc * 64 + 32 * t + l32;
        av[t] = r < m;
        bv[t] = c < n;
        ar[t] = it.A + (size_t)(av[t] ? r : 0) * it.lda;
        br[t] = bt + (size_t)(bv[t] ? c : 0) * it.ldbt;
    }
    auto frag = [&](const uint64_t *row, bool ok, int s) -> uint4 {
        const int w = 4 * s + 2 * half;
        const uint64_t x = ok && w < kw ? row[w] : 0ull;
        const uint64_t y = ok && w + 1 < kw ? row[w + 1] : 0ull;
        return make_uint4((uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32));
    };
    pk::v16f acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[x][y][r] = 8388608.0f;
    uint4 fa[2], fb[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        fa[t] = frag(ar[t], av[t], 0);
        fb[t] = frag(br[t], bv[t], 0);
    }
    for (int s = 0; s < nst; ++s) {
        const uint4 ga[2] = {fa[0], fa[1]}, gb[2] = {fb[0], fb[1]};
        if (s + 1 < nst)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                fa[t] = frag(ar[t], av[t], s + 1);
                fb[t] = frag(br[t], bv[t], s + 1);
            }
        pk::mm<0>(ga, gb, acc);
        pk::mm<1>(ga, gb, acc);
        pk::mm<2>(ga, gb, acc);
        pk::mm<3>(ga, gb, acc);
    }
    pk::store_parity_block(acc[0][0], acc[0][1], acc[1][0], acc[1][1], row0 + wr * 64, col0 + wc * 64, it.C,
                           it.ldc, m, n);
}

// Queue tiles a workgroup takes per atomic: a row-XOR tile is a few row
// loads per wave, less than the take itself (atomic, two barriers, the item
// search), so takes come in runs of consecutive tiles.
constexpr int MM_TAKE = 8;

// The persistent product: every workgroup takes runs of queue tiles until
// none is left (the tile index and the item are workgroup-uniform).  The
// item of a run's first tile is found by binary search; later tiles of the
// run step to the next item when they pass its first tile.
__global__ void __launch_bounds__(256)
k_mm_batch(const MMItem *__restrict__ items, int nitems, uint64_t total_tiles,
           unsigned long long *__restrict__ counter, const uint64_t *__restrict__ arena) {
    __shared__ unsigned long long s_t;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(counter, (unsigned long long)MM_TAKE);
        __syncthreads();
        const uint64_t t0 = s_t;
        __syncthreads();  // every thread has read s_t before the next take
        if (t0 >= total_tiles) break;
        const uint64_t t1 = t0 + MM_TAKE < total_tiles ? t0 + MM_TAKE : total_tiles;
        int ii = find_item(nitems, t0, [&](int j) { return items[j].tile_begin; });
        for (uint64_t t = t0; t < t1; ++t) {
        while (ii + 1 < nitems && items[ii + 1].tile_begin <= t) ++ii;
        const MMItem it = items[ii];
        const uint64_t lt = t - it.tile_begin;
        const uint32_t tx = (uint32_t)(lt % it.tiles_x), ty = (uint32_t)(lt / it.tiles_x);
        if (it.algo == MM_ANDPOP) {
            dev::andpop_tile(it.A, it.lda, arena + it.bt_off, it.ldbt, it.C, it.ldc, (int)it.m, (int)it.n,
                             (int)((it.k + 31) / 32), (int)(tx * AP_ROWS), (int)(ty * AP_COLS));
        } else if (it.algo == MM_MFMA) {
            mfma_tile(it, arena + it.bt_off, (int)(tx * MF_TILE), (int)(ty * MF_TILE));
        } else {
            const int row = (int)(tx * RX_WAVES) + wave;
            if (row < (int)it.m)
                dev::rowxor_unit(it.A, it.lda, it.B, it.ldb, it.C, it.ldc, row, (int)(ty * 64) + lane,
                                 (int)((it.k + 63) / 64), (int)((it.n + 63) / 64));
        }
        }
    }
}

// ones[i] = set bits of A_i (one workgroup per item)
__global__ void __launch_bounds__(256)
k_mm_popcount(const MMItem *__restrict__ items, unsigned long long *__restrict__ ones) {
    const MMItem it = items[blockIdx.x];
    const uint32_t kw = (it.k + 63) / 64;
    unsigned long long acc = 0;
    for (size_t e = threadIdx.x; e < (size_t)it.m * kw; e += 256) {
        const size_t r = e / kw, q = e - r * kw;
        acc += (unsigned long long)__popcll(it.A[r * it.lda + q]);
    }
    __shared__ unsigned long long part[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) ones[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

}  // namespace
}  // namespace mela

struct mela_matmul_batch {
    mela_ctx *ctx = nullptr;
    int flags = 0;
    std::vector<mela::MMItem> items;   // in caller order (algo chosen per item)
    std::vector<unsigned long long> ones;
    mela::MMItem *d_items = nullptr;   // in queue order
    unsigned long long *d_counter = nullptr;
    uint64_t *d_arena = nullptr;
    size_t arena_words = 0;
    uint64_t total_tiles = 0, total_tblocks = 0;
    int grid = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

using namespace mela;

namespace {

// Estimated seconds of one item under each engine (rates of the kernels on
// MI355X: AND+parity ~1.5e15 bit-op/s over whole 64 x 256 tiles, the
// matrix-core tiles MF_RATE over whole 128 x 128 tiles and 256-bit stages,
// row XOR ~1.6e13 bytes/s of B rows, whole 64-word chunks).
constexpr double MF_RATE = 8.0e15;
double est_andpop(const MMItem &it) {
    const double mp = (double)((it.m + AP_ROWS - 1) / AP_ROWS * AP_ROWS);
    const double np = (double)((it.n + AP_COLS - 1) / AP_COLS * AP_COLS);
    return 2.0 * mp * np * (double)it.k / 1.5e15;
}
// the matrix-core tiles: 128 x 128, k in whole 256-bit stages
double est_mfma(const MMItem &it) {
    const double mp = (double)((it.m + MF_TILE - 1) / MF_TILE * MF_TILE);
    const double np = (double)((it.n + MF_TILE - 1) / MF_TILE * MF_TILE);
    const double kp = (double)((it.k + 255) / 256 * 256);
    return 2.0 * mp * np * kp / MF_RATE;
}
double est_rowxor(const MMItem &it, unsigned long long ones) {
    const double row_bytes = (double)std::max<uint32_t>((it.n + 63) / 64, 64) * 8.0;
    return (double)ones * row_bytes / 1.6e13 + (double)it.m * ((it.k + 4095) / 4096) * 1e-10;
}

void drop_graph(mela_matmul_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void free_mm_batch(mela_matmul_batch *b) {
    drop_graph(b);
    for (void *p : {(void *)b->d_items, (void *)b->d_counter, (void *)b->d_arena})
        if (p) (void)hipFree(p);
    b->d_items = nullptr;
    b->d_counter = nullptr;
    b->d_arena = nullptr;
}

// Queue order, tiles, transpose blocks and the B^T arena for the current
// engine assignment; uploads the queue (the caller holds the context lock).
mela_status plan(mela_matmul_batch *b) {
    drop_graph(b);
    const size_t nb = b->items.size();
    std::vector<double> cost(nb);
    for (size_t i = 0; i < nb; ++i) {
        MMItem &it = b->items[i];
        if (it.k == 0) it.algo = MM_ROWXOR;  // C = 0: the row-XOR unit writes zeros
        if (it.algo == MM_ANDPOP) {
            it.tiles_x = (it.m + AP_ROWS - 1) / AP_ROWS;
            cost[i] = est_andpop(it);
        } else if (it.algo == MM_MFMA) {
            it.tiles_x = (it.m + MF_TILE - 1) / MF_TILE;
            cost[i] = est_mfma(it);
        } else {
            it.tiles_x = (it.m + RX_WAVES - 1) / RX_WAVES;
            cost[i] = est_rowxor(it, b->ones[i]);
        }
    }
    std::vector<int> order(nb);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    std::vector<MMItem> q(nb);
    uint64_t tiles = 0, tblocks = 0;
    size_t arena = 0;
    for (size_t p = 0; p < nb; ++p) {
        MMItem it = b->items[(size_t)order[p]];
        it.tile_begin = tiles;
        it.tb_begin = tblocks;
        if (it.m && it.n) {
            const uint64_t ty = it.algo == MM_ANDPOP ? (it.n + AP_COLS - 1) / AP_COLS
                                : it.algo == MM_MFMA ? (it.n + MF_TILE - 1) / MF_TILE
                                                     : ((it.n + 63) / 64 + 63) / 64;
            tiles += (uint64_t)it.tiles_x * ty;
        }
        if ((it.algo == MM_ANDPOP || it.algo == MM_MFMA) && it.m && it.n) {
            it.ldbt = (it.k + 63) / 64;
            it.bt_off = arena;
            arena += (size_t)it.n * it.ldbt;
            tblocks += (uint64_t)((it.k + 63) / 64) * (((it.n + 63) / 64 + MT_WORDS - 1) / MT_WORDS);
        }
        q[p] = it;
    }
    b->total_tiles = tiles;
    b->total_tblocks = tblocks;
    mela_ctx *ctx = b->ctx;
    if (arena > b->arena_words || !b->d_arena) {
        if (b->d_arena) (void)hipFree(b->d_arena);
        b->d_arena = nullptr;
        hipError_t e = hipMalloc((void **)&b->d_arena, std::max<size_t>(arena, 1) * 8);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            b->d_arena = nullptr;
            b->arena_words = 0;
            return set_error(MELA_ERR_NOMEM, "Out of memory.", "mela:memory");
        }
        b->arena_words = std::max<size_t>(arena, 1);
    }
    if (nb) MELA_HIP(hipMemcpyAsync(b->d_items, q.data(), nb * sizeof(MMItem), hipMemcpyHostToDevice, ctx->stream));
    MELA_HIP(hipStreamSynchronize(ctx->stream));
    return MELA_OK;
}

mela_status enqueue(mela_matmul_batch *b, hipStream_t s) {
    const int nb = (int)b->items.size();
    hipLaunchKernelGGL(k_mm_reset, dim3(1), dim3(64), 0, s, b->d_counter);
    if (b->total_tblocks) {
        const uint64_t wgs = (b->total_tblocks + 3) / 4;
        const unsigned g = (unsigned)std::min<uint64_t>(wgs, (uint64_t)b->ctx->cu_count * 8);
        hipLaunchKernelGGL(k_mm_transpose, dim3(g), dim3(256), 0, s, b->d_items, nb, b->total_tblocks,
                           b->d_arena);
    }
    if (b->total_tiles)
        hipLaunchKernelGGL(k_mm_batch, dim3((unsigned)b->grid), dim3(256), 0, s, b->d_items, nb,
                           b->total_tiles, b->d_counter, (const uint64_t *)b->d_arena);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

mela_status check_mat(const mela_bitmat &M) {
    if ((M.rows && M.cols && !M.data) || (M.cols && M.ld < words64(M.cols)))
        return set_error(MELA_ERR_INVALID, "bad matrix descriptor in batch");
    if (M.rows > 0xffffffffull || M.cols > 0xffffffffull || M.ld > 0xffffffffull)
        return set_error(MELA_ERR_INVALID, "batch matrices are limited to 2^32-1 rows and columns");
    return MELA_OK;
}

}  // namespace

extern "C" {

mela_status mela_matmul_batch_create(mela_ctx *ctx, const mela_bitmat *A, const mela_bitmat *B,
                                     const mela_bitmat *C, size_t count, int flags,
                                     mela_matmul_batch **out) {
    if (!ctx || !out || (count && (!A || !B || !C))) return set_error(MELA_ERR_INVALID, "null argument");
    *out = nullptr;
    std::vector<MMItem> items(count);
    for (size_t i = 0; i < count; ++i) {
        mela_status st;
        if ((st = check_mat(A[i])) || (st = check_mat(B[i])) || (st = check_mat(C[i]))) return st;
        if (A[i].cols != B[i].rows) return set_error(MELA_ERR_DIM, "Inner dimensions must agree.");
        if (C[i].rows != A[i].rows || C[i].cols != B[i].cols)
            return set_error(MELA_ERR_DIM, "output shape must be A.rows x B.cols");
        MMItem &it = items[i];
        it = MMItem{};
        it.A = A[i].data;
        it.B = B[i].data;
        it.C = C[i].data;
        it.m = (uint32_t)A[i].rows;
        it.k = (uint32_t)A[i].cols;
        it.n = (uint32_t)B[i].cols;
        it.lda = (uint32_t)A[i].ld;
        it.ldb = (uint32_t)B[i].ld;
        it.ldc = (uint32_t)C[i].ld;
    }
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    mela_matmul_batch *b = new mela_matmul_batch();
    b->ctx = ctx;
    b->flags = flags;
    b->items = std::move(items);
    const size_t nb = b->items.size();
    hipStream_t s = ctx->stream;
    auto fail = [&](mela_status st) {
        free_mm_batch(b);
        delete b;
        return st;
    };
    if (hipMalloc((void **)&b->d_items, std::max<size_t>(nb, 1) * sizeof(MMItem)) != hipSuccess ||
        hipMalloc((void **)&b->d_counter, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_error(MELA_ERR_NOMEM, "Out of memory.", "mela:memory"));
    }
    // density of every A_i, then the engine of each item
    b->ones.assign(nb, 0ull);
    if (nb) {
        unsigned long long *d_ones = nullptr;
        if (hipMalloc((void **)&d_ones, nb * sizeof(unsigned long long)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(set_error(MELA_ERR_NOMEM, "Out of memory.", "mela:memory"));
        }
        hipError_t e = hipMemcpyAsync(b->d_items, b->items.data(), nb * sizeof(MMItem), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_mm_popcount, dim3((unsigned)nb), dim3(256), 0, s, b->d_items, d_ones);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->ones.data(), d_ones, nb * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(d_ones);
        if (e != hipSuccess) return fail(hip_fail(e, "batch density count"));
    }
    for (size_t i = 0; i < nb; ++i) {
        MMItem &it = b->items[i];
        // the matrix cores beat AND+parity on every item shape measured
        // (forcing the planner's AND+parity items onto them only gained), so
        // the choice is theirs (AND+parity past their k limit) or row XOR
        const bool mf = it.k < MF_KMAX;
        const double ed = mf ? est_mfma(it) : est_andpop(it);
        it.algo = est_rowxor(it, b->ones[i]) < ed ? MM_ROWXOR : mf ? MM_MFMA : MM_ANDPOP;
    }
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)k_mm_batch, 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        occ = 1;
    }
    b->grid = std::max(1, occ) * ctx->cu_count;
    mela_status st = plan(b);
    if (st != MELA_OK) return fail(st);
    *out = b;
    return MELA_OK;
}

mela_status mela_matmul_batch_set_algos(mela_matmul_batch *b, const int *algos) {
    if (!b || (!b->items.empty() && !algos)) return set_error(MELA_ERR_INVALID, "null argument");
    for (size_t i = 0; i < b->items.size(); ++i) {
        if (algos[i] != MELA_ALGO_ANDPOP && algos[i] != MELA_ALGO_ROWXOR && algos[i] != MELA_ALGO_MFMA_FP4)
            return set_error(MELA_ERR_INVALID,
                             "batched products take MELA_ALGO_ANDPOP, MELA_ALGO_ROWXOR or MELA_ALGO_MFMA_FP4");
        if (algos[i] == MELA_ALGO_MFMA_FP4 && b->items[i].k >= MF_KMAX)
            return set_error(MELA_ERR_INVALID, "batched matrix-core products need k < 2^23");
    }
    CtxLock lk(b->ctx->mu);
    MELA_HIP(hipSetDevice(b->ctx->device));
    for (size_t i = 0; i < b->items.size(); ++i)
        b->items[i].algo = algos[i] == MELA_ALGO_ROWXOR ? MM_ROWXOR : algos[i] == MELA_ALGO_MFMA_FP4 ? MM_MFMA : MM_ANDPOP;
    return plan(b);
}

mela_status mela_matmul_batch_get_algos(mela_matmul_batch *b, int *algos) {
    if (!b || (!b->items.empty() && !algos)) return set_error(MELA_ERR_INVALID, "null argument");
    for (size_t i = 0; i < b->items.size(); ++i)
        algos[i] = b->items[i].algo == MM_ROWXOR ? MELA_ALGO_ROWXOR
                   : b->items[i].algo == MM_MFMA ? MELA_ALGO_MFMA_FP4
                                                 : MELA_ALGO_ANDPOP;
    return MELA_OK;
}

mela_status mela_matmul_batch_run(mela_matmul_batch *b) {
    if (!b) return set_error(MELA_ERR_INVALID, "null batch");
    mela_ctx *ctx = b->ctx;
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if (b->items.empty()) return MELA_OK;
    if (b->flags & MELA_BATCH_GRAPH) {
        if (!b->exec) {
            MELA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            mela_status st = enqueue(b, s);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            if (st != MELA_OK) return st;
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            b->graph = g;
            MELA_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        }
        MELA_HIP(hipGraphLaunch(b->exec, s));
        return MELA_OK;
    }
    return enqueue(b, s);
}

void mela_matmul_batch_destroy(mela_matmul_batch *b) {
    if (!b) return;
    {
        CtxLock lk(b->ctx->mu);
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        free_mm_batch(b);
    }
    delete b;
}

}  // extern "C"
