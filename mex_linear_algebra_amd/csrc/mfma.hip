// Exact GF(2) product on the matrix cores.
//
// A (m x k bits) and B^T (n x k bits) are expanded once to 4-bit e2m1 floats
// (bit 1 -> 1.0 = 0b0010) or to int8 (bit 1 -> 1); the product is then an
// ordinary matrix multiply: fp4 operands with fp32 accumulation
// (v_mfma_scale_f32_32x32x64_f8f6f4, unit scales) or int8 with int32
// accumulation (v_mfma_i32_32x32x32_i8).  Every partial sum is an integer
// <= k < 2^24, so both are exact, and C(i,j) = acc & 1 (the reference's
// A*B mod 2, matmul_gf2.c:110-208).
//
// Both instructions take 16 bytes per lane per operand (32 fp4 or 16 int8
// values of k), so one kernel template serves both formats.  A and B
// fragments use the same lane/byte -> k assignment, which makes the k order
// inside a fragment irrelevant to the result.
#include <type_traits>

#include "common.h"

namespace mela {
namespace {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int MF_UNITS = 4;    // 16-byte units of k per row per LDS stage (64 bytes)

// bits -> expanded bytes.  FP4: each source bit becomes one nibble (0x0 or
// 0x2), two per byte (element 2i low nibble); INT8: one byte (0 or 1).
template <bool FP4>
__global__ void __launch_bounds__(256)
k_expand(const uint64_t *__restrict__ src, size_t lds, size_t rows, size_t kwords,
         uint8_t *__restrict__ dst, size_t ldd_bytes) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t i = idx / kwords, w = idx % kwords;
    if (i >= rows) return;
    const uint64_t x = src[i * lds + w];
    if (FP4) {
        // 64 bits -> 32 bytes
        uint4 out[2];
        uint32_t *o = reinterpret_cast<uint32_t *>(out);
#pragma unroll
        for (int q = 0; q < 8; ++q) {  // 8 bits -> one dword (4 bytes)
            const uint32_t b = (uint32_t)(x >> (8 * q)) & 0xffu;
            uint32_t d = 0;
#pragma unroll
            for (int t = 0; t < 8; ++t) d |= ((b >> t) & 1u) << (4 * t + 1);
            o[q] = d;
        }
        uint4 *p = reinterpret_cast<uint4 *>(dst + i * ldd_bytes + w * 32);
        p[0] = out[0];
        p[1] = out[1];
    } else {
        // 64 bits -> 64 bytes
        uint4 out[4];
        uint32_t *o = reinterpret_cast<uint32_t *>(out);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t b = (uint32_t)(x >> (4 * q)) & 0xfu;
            o[q] = (b & 1u) | ((b & 2u) << 7) | ((b & 4u) << 14) | ((b & 8u) << 21);
        }
        uint4 *p = reinterpret_cast<uint4 *>(dst + i * ldd_bytes + w * 64);
#pragma unroll
        for (int q = 0; q < 4; ++q) p[q] = out[q];
    }
}

// C tile (2*WT) x (2*WT) per workgroup; 4 waves in a 2 x 2 grid, each WT x WT
// (WT/32 x WT/32 MFMA 32x32 tiles, accumulators in AGPRs).  k advances 64
// bytes of expanded data per LDS stage (two MFMA k-blocks of 32 bytes),
// double buffered.  LDS is stored fragment-major ([unit][row]) so each
// lane's 16-byte fragment read is contiguous across the wave.  WT = 128 gives
// 256 x 256 tiles: 32 KiB of operands per stage for 128 MFMAs.
template <bool FP4, int WT>
__global__ void __launch_bounds__(256)
k_mfma_gf2(const uint8_t *__restrict__ A, size_t lda, const uint8_t *__restrict__ B,
           size_t ldb, uint64_t *__restrict__ C, size_t ldc, int m, int n, int kbytes) {
    constexpr int TR = WT / 32;          // MFMA tiles per wave, each direction
    constexpr int ROWS = 2 * WT;         // workgroup tile rows (= columns)
    constexpr int LPT = ROWS * MF_UNITS / 256;  // 16-byte chunks per thread per operand
    __shared__ uint4 sA[2][MF_UNITS][ROWS];
    __shared__ uint4 sB[2][MF_UNITS][ROWS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    // row tiles walk fastest so resident workgroups share B tiles in L2
    const int i0 = blockIdx.x * ROWS, j0 = blockIdx.y * ROWS;

    // loader: thread t takes row t % ROWS of units t / ROWS + q * (256 / ROWS);
    // consecutive threads read consecutive rows (each 16 bytes of one unit).
    // Named registers (no array) keep the staged chunks out of scratch.
    const int lrow = tid % ROWS, lu0 = tid / ROWS;
    constexpr int USTEP = 256 / ROWS;
    const bool a_ok = i0 + lrow < m, b_ok = j0 + lrow < n;
    const uint8_t *a_src = A + (size_t)(a_ok ? i0 + lrow : 0) * lda;
    const uint8_t *b_src = B + (size_t)(b_ok ? j0 + lrow : 0) * ldb;
    uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
    const uint4 zero4 = make_uint4(0, 0, 0, 0);
#define MF_LD1(q, RA, RB)                                                                       \
    do {                                                                                        \
        const int off = kb0 + (lu0 + (q) * USTEP) * 16;                                         \
        const bool in = off < kbytes;                                                           \
        RA = (a_ok && in) ? *reinterpret_cast<const uint4 *>(a_src + off) : zero4;              \
        RB = (b_ok && in) ? *reinterpret_cast<const uint4 *>(b_src + off) : zero4;              \
    } while (0)
    auto gload = [&](int kb0) __attribute__((always_inline)) {
        MF_LD1(0, ra0, rb0);
        if constexpr (LPT > 1) MF_LD1(1, ra1, rb1);
        if constexpr (LPT > 2) MF_LD1(2, ra2, rb2);
        if constexpr (LPT > 3) MF_LD1(3, ra3, rb3);
    };
#undef MF_LD1
    auto sstore = [&](int buf) __attribute__((always_inline)) {
        sA[buf][lu0][lrow] = ra0;
        sB[buf][lu0][lrow] = rb0;
        if constexpr (LPT > 1) {
            sA[buf][lu0 + USTEP][lrow] = ra1;
            sB[buf][lu0 + USTEP][lrow] = rb1;
        }
        if constexpr (LPT > 2) {
            sA[buf][lu0 + 2 * USTEP][lrow] = ra2;
            sB[buf][lu0 + 2 * USTEP][lrow] = rb2;
        }
        if constexpr (LPT > 3) {
            sA[buf][lu0 + 3 * USTEP][lrow] = ra3;
            sB[buf][lu0 + 3 * USTEP][lrow] = rb3;
        }
    };

    using Acc = typename std::conditional<FP4, v16f, v16i>::type;
    Acc acc[TR][TR];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;

    const int nstage = (kbytes + MF_UNITS * 16 - 1) / (MF_UNITS * 16);
    gload(0);
    sstore(0);
    __syncthreads();
    const int half = lane >> 5, l32 = lane & 31;
    for (int st = 0; st < nstage; ++st) {
        const int buf = st & 1;
        if (st + 1 < nstage) gload((st + 1) * MF_UNITS * 16);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            uint4 fa[TR], fb[TR];
#pragma unroll
            for (int t = 0; t < TR; ++t) {
                fa[t] = sA[buf][kb * 2 + half][wr * WT + t * 32 + l32];
                fb[t] = sB[buf][kb * 2 + half][wc * WT + t * 32 + l32];
            }
#pragma unroll
            for (int tr = 0; tr < TR; ++tr)
#pragma unroll
                for (int tc = 0; tc < TR; ++tc) {
                    if constexpr (FP4) {
                        v8i a8 = {(int)fa[tr].x, (int)fa[tr].y, (int)fa[tr].z, (int)fa[tr].w, 0, 0, 0, 0};
                        v8i b8 = {(int)fb[tc].x, (int)fb[tc].y, (int)fb[tc].z, (int)fb[tc].w, 0, 0, 0, 0};
                        acc[tr][tc] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                            a8, b8, acc[tr][tc], 4, 4, 0, 127, 0, 127);
                    } else {
                        v4i a4 = {(int)fa[tr].x, (int)fa[tr].y, (int)fa[tr].z, (int)fa[tr].w};
                        v4i b4 = {(int)fb[tc].x, (int)fb[tc].y, (int)fb[tc].z, (int)fb[tc].w};
                        acc[tr][tc] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a4, b4, acc[tr][tc], 0, 0, 0);
                    }
                }
        }
        if (st + 1 < nstage) sstore(buf ^ 1);
        __syncthreads();
    }

    // 32x32 accumulator layout: register r of lane l holds row
    // 8*(r/4) + 4*(l/32) + r%4 and column l%32 of its tile.  Column tiles
    // (2c, 2c+1) of a wave form one 64-bit word per row.
    const int nw = (n + 63) >> 6;
#pragma unroll
    for (int tr = 0; tr < TR; ++tr)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int rbase = i0 + wr * WT + tr * 32 + 8 * (r >> 2) + (r & 3);
#pragma unroll
            for (int tc = 0; tc < TR; tc += 2) {
                const uint64_t b0 = __ballot((int)acc[tr][tc][r] & 1);
                const uint64_t b1 = __ballot((int)acc[tr][tc + 1][r] & 1);
                const int w = (j0 + wc * WT + tc * 32) >> 6;
                // lane 0 writes the row of lanes 0-31, lane 32 that of lanes 32-63
                if ((lane & 31) == 0 && w < nw) {
                    const int sh = lane >> 5;
                    const int row = rbase + 4 * sh;
                    const uint64_t lo = sh ? (b0 >> 32) : (b0 & 0xffffffffull);
                    const uint64_t hi = sh ? (b1 >> 32) : (b1 & 0xffffffffull);
                    if (row < m) C[(size_t)row * ldc + w] = lo | (hi << 32);
                }
            }
        }
}

}  // namespace

// C = A * B via the matrix cores; fp4 (fmt 4) or int8 expansion.
mela_status matmul_mfma(mela_ctx *ctx, const uint64_t *A, size_t lda, const uint64_t *B,
                        size_t ldb, uint64_t *C, size_t ldc, size_t m, size_t k, size_t n,
                        bool fp4) {
    hipStream_t s = ctx->stream;
    const size_t kw = words64(k);
    const size_t row_bytes = kw * (fp4 ? 32 : 64);  // expanded bytes per row
    void *pa = nullptr, *pb = nullptr, *pbt = nullptr;
    mela_status st;
    if ((st = scratch_get(ctx, SLOT_BT, n * kw * 8, &pbt)) != MELA_OK) return st;
    if ((st = scratch_get(ctx, SLOT_XA, m * row_bytes, &pa)) != MELA_OK) return st;
    if ((st = scratch_get(ctx, SLOT_XB, n * row_bytes, &pb)) != MELA_OK) return st;
    if ((st = launch_transpose(s, B, ldb, k, n, (uint64_t *)pbt, kw)) != MELA_OK) return st;
    const size_t ta = m * kw, tb = n * kw;
    if (fp4) {
        hipLaunchKernelGGL(k_expand<true>, dim3((unsigned)((ta + 255) / 256)), dim3(256), 0, s, A,
                           lda, m, kw, (uint8_t *)pa, row_bytes);
        hipLaunchKernelGGL(k_expand<true>, dim3((unsigned)((tb + 255) / 256)), dim3(256), 0, s,
                           (const uint64_t *)pbt, kw, n, kw, (uint8_t *)pb, row_bytes);
    } else {
        hipLaunchKernelGGL(k_expand<false>, dim3((unsigned)((ta + 255) / 256)), dim3(256), 0, s, A,
                           lda, m, kw, (uint8_t *)pa, row_bytes);
        hipLaunchKernelGGL(k_expand<false>, dim3((unsigned)((tb + 255) / 256)), dim3(256), 0, s,
                           (const uint64_t *)pbt, kw, n, kw, (uint8_t *)pb, row_bytes);
    }
    // 128 x 128 tiles (256 x 256 tiles measured slower: one wave per SIMD
    // leaves the global-load latency of each stage exposed)
    constexpr unsigned T = 128;
    dim3 grid((unsigned)((m + T - 1) / T), (unsigned)((n + T - 1) / T));
    if (grid.y > 65535) return set_error(MELA_ERR_INVALID, "too many columns for one launch");
    const uint8_t *xa = (const uint8_t *)pa, *xb = (const uint8_t *)pb;
    const int kb = (int)row_bytes;
    if (fp4)
        hipLaunchKernelGGL((k_mfma_gf2<true, 64>), grid, dim3(256), 0, s, xa, row_bytes, xb, row_bytes, C, ldc, (int)m, (int)n, kb);
    else
        hipLaunchKernelGGL((k_mfma_gf2<false, 64>), grid, dim3(256), 0, s, xa, row_bytes, xb, row_bytes, C, ldc, (int)m, (int)n, kb);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

}  // namespace mela
