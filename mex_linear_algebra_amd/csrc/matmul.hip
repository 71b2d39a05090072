// GF(2) matrix product C = A * B on packed operands.
//
// Two exact kernels, both bit-identical to the reference cores
// (matmul_gf2.c:110-208 and matmul_m4ri.c:178-248, whose result is A*B mod 2):
//
//  * k_andpop  -- C(i,j) = parity(popcount(A_i AND Bt_j)).  Per 32 k-bits one
//    v_bitop3_b32 does the AND and the XOR-accumulate; the popcount parity is
//    taken once per output bit at the end and the 64 bits of a word are
//    assembled by a wave ballot.  Bound: VALU bitwise issue rate.
//  * k_m4rm    -- Method of Four Russians.  For every 8-row strip of B the
//    workgroup builds the 256 XOR-combinations of those rows (for its 2048
//    output columns) in LDS; each output row then XORs one table row selected
//    by the 8 matching bits of A.  Bound: LDS read bandwidth / VALU issue.
#include <algorithm>

#include "common.h"
#include "product_device.h"

namespace mela {
namespace {

using dev::AP_ROWS;
using dev::AP_COLS;
using dev::RX_WAVES;

// ------------------------------------------------------------ AND+parity
// (tile function in product_device.h)
// blockIdx.x walks row tiles fastest: the workgroups resident together on an
// XCD share one B^T column tile (1 MiB at k = 32768) in that XCD's L2.
__global__ void __launch_bounds__(256)
k_andpop(const uint64_t *__restrict__ A, size_t lda, const uint64_t *__restrict__ Bt,
         size_t ldb, uint64_t *__restrict__ C, size_t ldc, int m, int n, int kd) {
    dev::andpop_tile(A, lda, Bt, ldb, C, ldc, m, n, kd, blockIdx.x * AP_ROWS, blockIdx.y * AP_COLS);
}

// --------------------------------------------------------------- row XOR
// (unit function in product_device.h)
// The grid is at most a few workgroups per CU, each striding over the (row,
// chunk) pairs: when the density gate turns the kernel off, thousands of
// empty workgroups would cost more dispatch time than the product.
__global__ void __launch_bounds__(RX_WAVES * 64)
k_rowxor(const uint64_t *__restrict__ A, size_t lda, const uint64_t *__restrict__ B, size_t ldb,
         uint64_t *__restrict__ C, size_t ldc, int m, int kw, int nw, Gate gate) {
    if (gate_skip(gate, true)) return;
    const int lane = threadIdx.x & 63;
    const int nch = (nw + 63) / 64;
    const long units = (long)m * nch;  // (row, 64-word chunk) pairs
    for (long u = __builtin_amdgcn_readfirstlane((int)blockIdx.x * RX_WAVES + (threadIdx.x >> 6)); u < units;
         u += (long)gridDim.x * RX_WAVES)
        dev::rowxor_unit(A, lda, B, ldb, C, ldc, (int)(u / nch), (int)(u % nch) * 64 + lane, kw, nw);
}

// ones[b % GATE_SLOTS] += popcount of workgroup b's share of the rows x kw
// words of A (the density of the AUTO choice)
__global__ void __launch_bounds__(256)
k_popcount_rows(const uint64_t *__restrict__ A, size_t lda, size_t rows, size_t kw,
                unsigned long long *__restrict__ ones) {
    unsigned long long acc = 0;
    const size_t tot = rows * kw;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < tot; e += (size_t)gridDim.x * 256) {
        const size_t r = e / kw, q = e - r * kw;
        acc += (unsigned long long)__popcll(A[r * lda + q]);
    }
    // wave sums, then one atomic per workgroup (the grid is a few hundred
    // workgroups: one atomic per wave on a single address cost ~80 us)
    __shared__ unsigned long long part[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(ones + blockIdx.x % GATE_SLOTS, part[0] + part[1] + part[2] + part[3]);
}

// ------------------------------------------------------------------ M4RM
constexpr int M4_WAVES = 16;
constexpr int M4_ROWS_PER_WAVE = 64;
constexpr int M4_ROWS = M4_WAVES * M4_ROWS_PER_WAVE;  // 1024 rows per workgroup
constexpr int M4_DW = 64;                             // dwords (2048 columns) per workgroup

// Builds this wave's 16 entries (16*wave .. 16*wave+15) of the 256-entry table
// of XOR-combinations of the 8 B rows held (for this lane's dword) in br[]:
// the high nibble (= wave) selects rows 4..7 for the base, then a Gray-code
// walk over the low nibble changes one row per entry (one XOR, one live temp).
__device__ __forceinline__ void m4_build(uint32_t (*T)[M4_DW], const uint32_t br[8], int wave,
                                         int lane) {
    uint32_t cur = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if ((wave >> b) & 1) cur ^= br[4 + b];
    T[wave * 16][lane] = cur;
#pragma unroll
    for (int t = 1; t < 16; ++t) {
        cur ^= br[__builtin_ctz(t)];
        T[wave * 16 + (t ^ (t >> 1))][lane] = cur;
    }
}

// C tile: 1024 rows x 2048 columns; blockIdx.z selects a k-range [kb, ke) of
// 64-bit A words (split-K partial products are XOR-combined with atomics).
__global__ void __launch_bounds__(1024)
k_m4rm(const uint64_t *__restrict__ A, size_t lda, const uint64_t *__restrict__ B, size_t ldb,
       uint64_t *__restrict__ C, size_t ldc, int m, int n, int k, int kw_per_split, int atomic) {
    __shared__ uint32_t T[2][256][M4_DW];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int row0 = blockIdx.y * M4_ROWS + wave * M4_ROWS_PER_WAVE;
    const int dw = blockIdx.x * M4_DW + lane;  // this lane's output dword
    const int ndw = ((n + 63) >> 6) * 2;  // dwords of the packed row width
    const int kw_total = (k + 63) >> 6;
    const int kwb = blockIdx.z * kw_per_split;
    int kwe = kwb + kw_per_split;
    if (kwe > kw_total) kwe = kw_total;
    if (kwb >= kwe) return;
    const int s_begin = kwb * 8;
    const int s_end = min(kwe * 8, (k + 7) >> 3);  // 8-row strips

    const uint32_t *B32 = reinterpret_cast<const uint32_t *>(B);
    const size_t ldb32 = ldb * 2;
    const bool col_ok = dw < ndw;

    // lane r of this wave holds A row row0+r
    const int arow = row0 + lane;
    const bool arow_ok = arow < m;
    const uint64_t *a_src = A + (size_t)arow * lda;

    uint32_t acc[M4_ROWS_PER_WAVE];
#pragma unroll
    for (int r = 0; r < M4_ROWS_PER_WAVE; ++r) acc[r] = 0u;

    // branch-free strip loads: clamp the address into B, then mask the value
    const size_t dw_ld = col_ok ? (size_t)dw : (size_t)(ndw - 1);
    const uint32_t col_mask = col_ok ? ~0u : 0u;
    uint32_t br[8];
    auto load_strip = [&](int s) {
        const uint32_t *src = B32 + (size_t)s * 8 * ldb32 + dw_ld;
        if (s * 8 + 8 <= k) {
#pragma unroll
            for (int b = 0; b < 8; ++b) br[b] = src[(size_t)b * ldb32] & col_mask;
        } else {
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int kr = s * 8 + b;
                const int krc = kr < k ? kr : k - 1;
                const uint32_t v = B32[(size_t)krc * ldb32 + dw_ld];
                br[b] = (kr < k) ? (v & col_mask) : 0u;
            }
        }
    };

    load_strip(s_begin);
    m4_build(T[0], br, wave, lane);
    __syncthreads();

    uint64_t aw = arow_ok ? a_src[kwb] : 0ull;
    for (int s = s_begin; s < s_end; ++s) {
        const int buf = (s - s_begin) & 1;
        const bool more = s + 1 < s_end;
        if (more) load_strip(s + 1);
        const int sb = s & 7;
        const uint32_t half = (sb < 4) ? (uint32_t)aw : (uint32_t)(aw >> 32);
        const int sh = (sb & 3) * 8;
        // groups of 8 lookups; the scheduling barrier stops the compiler from
        // hoisting all 64 table addresses into registers at once
#pragma unroll
        for (int r0 = 0; r0 < M4_ROWS_PER_WAVE; r0 += 8) {
#pragma unroll
            for (int r = r0; r < r0 + 8; ++r) {
                const uint32_t ar = __builtin_amdgcn_readlane(half, r);
                const uint32_t idx = (ar >> sh) & 255u;
                acc[r] ^= T[buf][idx][lane];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (sb == 7 && more) aw = arow_ok ? a_src[(s + 1) >> 3] : 0ull;
        if (more) m4_build(T[buf ^ 1], br, wave, lane);
        __syncthreads();
    }

    if (!col_ok) return;
    uint32_t *C32 = reinterpret_cast<uint32_t *>(C);
    const size_t ldc32 = ldc * 2;
#pragma unroll
    for (int r = 0; r < M4_ROWS_PER_WAVE; ++r) {
        const int i = row0 + r;
        if (i < m) {
            if (atomic)
                atomicXor(&C32[(size_t)i * ldc32 + dw], acc[r]);
            else
                C32[(size_t)i * ldc32 + dw] = acc[r];
        }
    }
}

}  // namespace

mela_status launch_popcount(mela_ctx *ctx, const uint64_t *A, size_t lda, size_t rows, size_t kw,
                            unsigned long long *ones) {
    const size_t blocks = std::min<size_t>((rows * kw + 255) / 256, (size_t)ctx->cu_count * 2);
    hipLaunchKernelGGL(k_popcount_rows, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0,
                       ctx->stream, A, lda, rows, kw, ones);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

namespace {

// The AUTO product.  Dense kernels cost m*n*k bit operations whatever the
// data; the row-XOR product costs one B-row read per set bit of A.  Below
// 2^33 bit-products AUTO is AND+parity.  From there on it is the fp4 matrix
// cores, unless A is sparse: the bits of A are counted on the device (by the
// kernel that re-lays A for the matrix cores, or a separate pass) and both
// the row-XOR kernel and the remaining fp4 kernels are enqueued behind a
// gate (each exits at once unless the count picks it), so the choice costs
// no host round trip.  The threshold compares the
// measured rates on MI355X: fp4 ~7.5e15 bit-op/s plus ~15 us per call, row
// XOR ~1.6e13 bytes/s of B rows plus ~5 us.
Gate auto_gate(size_t m, size_t k, size_t n, unsigned long long *ones) {
    const double dense_s = 2.0 * (double)m * (double)n * (double)k / 7.5e15 + 15e-6;
    const double row_bytes = (double)std::max<size_t>(words64(n), 64) * 8.0;
    const double budget = dense_s - 5e-6 - (double)m * (double)((k + 4095) / 4096) * 1e-10;
    Gate g;
    g.ones = ones;
    g.thr = budget > 0 ? (unsigned long long)(budget * 1.6e13 / row_bytes) : 0ull;
    return g;
}

mela_status launch_rowxor(hipStream_t s, int cu_count, const uint64_t *A, size_t lda, const uint64_t *B,
                          size_t ldb, uint64_t *C, size_t ldc, size_t m, size_t k, size_t n, Gate gate) {
    const int kw = (int)words64(k), nwc = (int)words64(n);
    const size_t units = m * (size_t)((nwc + 63) / 64);
    const size_t grid = std::min<size_t>((units + RX_WAVES - 1) / RX_WAVES, (size_t)cu_count * 8);
    hipLaunchKernelGGL(k_rowxor, dim3((unsigned)std::max<size_t>(grid, 1)), dim3(RX_WAVES * 64), 0, s, A, lda,
                       B, ldb, C, ldc, (int)m, kw, nwc, gate);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

}  // namespace

mela_status matmul_device(mela_ctx *ctx, const uint64_t *A, size_t lda, const uint64_t *B,
                          size_t ldb, uint64_t *C, size_t ldc, size_t m, size_t k, size_t n,
                          int algo) {
    hipStream_t s = ctx->stream;
    const size_t nw = words64(n);
    if (m == 0 || n == 0) return MELA_OK;
    if (k == 0) return launch_fill_zero(s, C, m * ldc * sizeof(uint64_t));  // C = 0
    if (m > (size_t)INT32_MAX || n > (size_t)INT32_MAX || k > (size_t)INT32_MAX)
        return set_error(MELA_ERR_INVALID, "matrix dimension exceeds 2^31-1");
    if (algo == MELA_ALGO_AUTO) {
        const bool fp4 = (double)m * (double)n * (double)k >= 8589934592.0 && k < (1u << 24);
        if (!fp4) {
            algo = MELA_ALGO_ANDPOP;
        } else {
            void *p = nullptr;
            const size_t cb = GATE_SLOTS * sizeof(unsigned long long);
            mela_status st = scratch_get(ctx, SLOT_DENSITY, cb, &p);
            if (st != MELA_OK) return st;
            if ((st = launch_fill_zero(s, p, cb)) != MELA_OK) return st;
            unsigned long long *count = (unsigned long long *)p;
            const Gate gate = auto_gate(m, k, n, count);
            auto sparse = [&](Gate g) {
                return launch_rowxor(s, ctx->cu_count, A, lda, B, ldb, C, ldc, m, k, n, g);
            };
            return matmul_mfma(ctx, A, lda, B, ldb, C, ldc, m, k, n, true, gate, count, sparse);
        }
    }
    if (algo == MELA_ALGO_ROWXOR) {
        prof_mark(ctx, true);
        mela_status st = launch_rowxor(s, ctx->cu_count, A, lda, B, ldb, C, ldc, m, k, n, Gate());
        prof_mark(ctx, false);
        return st;
    }
    if (algo == MELA_ALGO_ANDPOP) {
        // Bt = B^T (n x ceil(k/64) words)
        const size_t ldbt = words64(k);
        void *bt = nullptr;
        mela_status st = scratch_get(ctx, SLOT_BT, n * ldbt * sizeof(uint64_t), &bt);
        if (st != MELA_OK) return st;
        st = launch_transpose(s, B, ldb, k, n, (uint64_t *)bt, ldbt);
        if (st != MELA_OK) return st;
        const int kd = (int)((k + 31) / 32);
        dim3 grid((unsigned)((m + AP_ROWS - 1) / AP_ROWS), (unsigned)((n + AP_COLS - 1) / AP_COLS));
        if (grid.y > 65535) return set_error(MELA_ERR_INVALID, "too many columns for one launch");
        prof_mark(ctx, true);
        hipLaunchKernelGGL(k_andpop, grid, dim3(256), 0, s, A, lda, (const uint64_t *)bt, ldbt, C,
                           ldc, (int)m, (int)n, kd);
        prof_mark(ctx, false);
        MELA_HIP(hipGetLastError());
        return MELA_OK;
    }
    if (algo == MELA_ALGO_M4RM) {
        const int kw = (int)words64(k);
        const unsigned gx = (unsigned)((n + 32 * M4_DW - 1) / (32 * M4_DW));
        const unsigned gy = (unsigned)((m + M4_ROWS - 1) / M4_ROWS);
        if (gy > 65535) return set_error(MELA_ERR_INVALID, "too many rows for one launch");
        // split k until the grid covers the chip twice (XOR partials are exact
        // and order-free, so the atomics cannot change the result)
        int splits = 1;
        const unsigned target = 2u * (unsigned)ctx->cu_count;
        while (gx * gy * splits < target && splits * 2 <= kw && kw / (splits * 2) >= 8) splits *= 2;
        const int per = (kw + splits - 1) / splits;
        int atomic = splits > 1;
        if (atomic) {
            mela_status st = launch_fill_zero(s, C, m * ldc * sizeof(uint64_t));
            if (st != MELA_OK) return st;
        }
        prof_mark(ctx, true);
        hipLaunchKernelGGL(k_m4rm, dim3(gx, gy, splits), dim3(1024), 0, s, A, lda, B, ldb, C, ldc,
                           (int)m, (int)n, (int)k, per, atomic);
        prof_mark(ctx, false);
        MELA_HIP(hipGetLastError());
        (void)nw;
        return MELA_OK;
    }
    if (algo == MELA_ALGO_MFMA_I8 || algo == MELA_ALGO_MFMA_FP4) {
        if (k >= (1u << 24))
            return set_error(MELA_ERR_INVALID, "MFMA path needs k < 2^24 (exact accumulation)");
        return matmul_mfma(ctx, A, lda, B, ldb, C, ldc, m, k, n, algo == MELA_ALGO_MFMA_FP4);
    }
    return set_error(MELA_ERR_INVALID, "unknown matmul algorithm");
}

}  // namespace mela
