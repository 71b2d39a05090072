// Workgroup-level GF(2) product tiles shared by the single-product kernels
// (matmul.hip) and the batched persistent product (batch_matmul.hip).
#pragma once

#include "common.h"

namespace mela {
namespace dev {

// ------------------------------------------------------------ AND+parity
constexpr int AP_MR = 16;   // rows per wave (all lanes share them)
constexpr int AP_NR = 4;    // 64-column slots per lane
constexpr int AP_KC = 16;   // k-dwords per LDS chunk
constexpr int AP_WAVES = 4;
constexpr int AP_ROWS = AP_MR * AP_WAVES;  // 64 rows per workgroup
constexpr int AP_COLS = 64 * AP_NR;        // 256 columns per workgroup

// A: m x lda words (packed along k), Bt: n x ldb words (B transposed, packed
// along k).  kd = ceil(k/32) dwords; bits past k are zero in both.
// One workgroup of 256 threads computes the AP_ROWS x AP_COLS tile of C at
// (i0, j0).  Called repeatedly by a persistent kernel, it leaves the LDS
// free: its last chunk ends with a barrier and the epilogue does not touch it.
__device__ __forceinline__ void andpop_tile(const uint64_t *__restrict__ A, size_t lda,
                                            const uint64_t *__restrict__ Bt, size_t ldb,
                                            uint64_t *__restrict__ C, size_t ldc, int m, int n,
                                            int kd, int i0, int j0) {
    __shared__ uint32_t sA[2][AP_ROWS][AP_KC];
    __shared__ uint32_t sB[2][AP_KC][64][AP_NR];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t *A32 = reinterpret_cast<const uint32_t *>(A);
    const uint32_t *B32 = reinterpret_cast<const uint32_t *>(Bt);
    const size_t lda32 = lda * 2, ldb32 = ldb * 2;

    // loader roles: A tile row = tid/4, 4 dwords at (tid%4)*4; B column = tid.
    const int la_row = tid >> 2, la_off = (tid & 3) * 4;
    const int lb_col = tid;
    const bool a_ok = i0 + la_row < m;
    const bool b_ok = j0 + lb_col < n;
    const uint32_t *a_src = A32 + (size_t)(i0 + la_row) * lda32;
    const uint32_t *b_src = B32 + (size_t)(j0 + lb_col) * ldb32;

    uint32_t ra[4], rb[AP_KC];
    auto gload = [&](int kd0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int d = kd0 + la_off + q;
            ra[q] = (a_ok && d < kd) ? a_src[d] : 0u;
        }
#pragma unroll
        for (int q = 0; q < AP_KC; ++q) {
            int d = kd0 + q;
            rb[q] = (b_ok && d < kd) ? b_src[d] : 0u;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) sA[buf][la_row][la_off + q] = ra[q];
#pragma unroll
        for (int q = 0; q < AP_KC; ++q) sB[buf][q][lb_col & 63][lb_col >> 6] = rb[q];
    };

    uint32_t acc[AP_MR][AP_NR];
#pragma unroll
    for (int r = 0; r < AP_MR; ++r)
#pragma unroll
        for (int c = 0; c < AP_NR; ++c) acc[r][c] = 0u;

    const int nchunks = (kd + AP_KC - 1) / AP_KC;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int ch = 0; ch < nchunks; ++ch) {
        const int buf = ch & 1;
        if (ch + 1 < nchunks) gload((ch + 1) * AP_KC);
#pragma unroll 1
        for (int kk = 0; kk < AP_KC; kk += 4) {
            uint4 bv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                bv[q] = *reinterpret_cast<const uint4 *>(&sB[buf][kk + q][lane][0]);
#pragma unroll
            for (int r = 0; r < AP_MR; ++r) {
                const uint4 av =
                    *reinterpret_cast<const uint4 *>(&sA[buf][wave * AP_MR + r][kk]);
                acc[r][0] ^= (av.x & bv[0].x) ^ (av.y & bv[1].x) ^ (av.z & bv[2].x) ^ (av.w & bv[3].x);
                acc[r][1] ^= (av.x & bv[0].y) ^ (av.y & bv[1].y) ^ (av.z & bv[2].y) ^ (av.w & bv[3].y);
                acc[r][2] ^= (av.x & bv[0].z) ^ (av.y & bv[1].z) ^ (av.z & bv[2].z) ^ (av.w & bv[3].z);
                acc[r][3] ^= (av.x & bv[0].w) ^ (av.y & bv[1].w) ^ (av.z & bv[2].w) ^ (av.w & bv[3].w);
            }
        }
        if (ch + 1 < nchunks) sstore(buf ^ 1);
        __syncthreads();
    }

    const int nw = (n + 63) >> 6;
#pragma unroll
    for (int r = 0; r < AP_MR; ++r) {
        const int i = i0 + wave * AP_MR + r;
#pragma unroll
        for (int c = 0; c < AP_NR; ++c) {
            uint64_t word = __ballot(__popc(acc[r][c]) & 1);
            const int w = (j0 >> 6) + c;
            if (lane == r * AP_NR + c && i < m && w < nw) C[(size_t)i * ldc + w] = word;
        }
    }
}

// --------------------------------------------------------------- row XOR
// C_i = XOR of the rows B_k over the set bits k of A_i: the reference's own
// GE product (matmul_gf2.c:127-186 XORs B_k into the row-block accumulator
// for every odd A(i,k)), whose cost is nnz(A) row reads instead of m*n*k bit
// operations, so it is the product of choice for sparse A.  One wave per
// (row i, 64-word chunk of the row of C): lanes are words.  A_i is read 64
// words per vector load (lane q holds word kb + q); each round takes the
// lowest set bit of up to RX_BATCH different nonzero words and has their
// rows of B in flight together, so a sparse row (about one set bit per
// nonzero word) costs one round trip per RX_BATCH bits, not one per bit.
// An empty slot re-reads the row of lane 0's lowest bit (or row 64 kb),
// which lies below k, and masks it.
constexpr int RX_WAVES = 4;
constexpr int RX_BATCH = 8;

// Row i of C, word w (lane-varying): XOR of B's word w over the set bits of
// A_i (i is wave-uniform; all 64 lanes of the wave take part).
__device__ __forceinline__ void rowxor_unit(const uint64_t *__restrict__ A, size_t lda,
                                            const uint64_t *__restrict__ B, size_t ldb,
                                            uint64_t *__restrict__ C, size_t ldc, int i, int w, int kw,
                                            int nw) {
    const int lane = threadIdx.x & 63;
    const bool wok = w < nw;
    const uint64_t *arow = A + (size_t)i * lda;
    const uint64_t *bcol = B + (wok ? w : 0);
    uint64_t c = 0;
    for (int kb = 0; kb < kw; kb += 64) {
        uint64_t x = kb + lane < kw ? arow[kb + lane] : 0ull;
        for (uint64_t nz = __ballot(x != 0ull); nz; nz = __ballot(x != 0ull)) {
            const int myb = x ? __builtin_ctzll(x) : 0;  // this word's lowest set bit
            uint64_t v[RX_BATCH], keep[RX_BATCH];
            uint64_t taken = 0;
#pragma unroll
            for (int q = 0; q < RX_BATCH; ++q) {
                const bool has = nz != 0ull;
                const int src = has ? __builtin_ctzll(nz) : 0;
                if (has) {
                    nz &= nz - 1ull;
                    taken |= 1ull << src;
                }
                const int bit = __builtin_amdgcn_readlane(myb, src);
                keep[q] = has ? ~0ull : 0ull;
                v[q] = bcol[(size_t)((kb + src) * 64 + bit) * ldb];
            }
#pragma unroll
            for (int q = 0; q < RX_BATCH; ++q) c ^= v[q] & keep[q];
            if ((taken >> lane) & 1ull) x &= x - 1ull;
        }
    }
    if (wok) C[(size_t)i * ldc + w] = c;
}

}  // namespace dev
}  // namespace mela
