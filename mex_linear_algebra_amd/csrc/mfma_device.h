// fp4 matrix-core steps on packed GF(2) bits, shared by the product kernel
// (mfma.hip) and the trailing update of the blocked elimination
// (elim_block.hip).
//
// Four planes per packed dword give the fp4 operands, by masking alone:
//   plane 0: x & 0x11111111        (bit 4t   -> nibble t, e2m1 0b0001 = 0.5)
//   plane 1: x & 0x22222222        (bit 4t+1 -> nibble t, 0b0010 = 1.0)
//   plane 2: x & 0x44444444        (bit 4t+2 -> nibble t, 0b0100 = 2.0)
//   plane 3: (x >> 1) & 0x44444444 (bit 4t+3 -> nibble t, 2.0)
// Each MFMA takes one plane of both operands, so every product in it is
// w * w for the plane's weight w, and the instruction's E8M0 block scales
// (2^1, 2^0, 2^-1, 2^-1 on each operand) make every product exactly 1.0: the
// fp32 accumulator counts the k with A(i,k) = B(k,j) = 1 exactly (< 2^24).
// A and B use the same map from (lane half, dword, nibble, plane) to k, so
// which k lands where inside a fragment does not matter.
#pragma once

#include "common.h"
#include "transpose_device.h"

namespace mela {
namespace pk {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Plane j of four packed dwords as an fp4 operand (the upper four dwords of
// the 8-dword operand are unused by fp4).
template <int J>
__device__ __forceinline__ v8i plane(const uint4 &x) {
    constexpr uint32_t M = J == 3 ? 0x44444444u : (0x11111111u << J);
    constexpr int SH = J == 3 ? 1 : 0;
    return v8i{(int)((x.x >> SH) & M), (int)((x.y >> SH) & M), (int)((x.z >> SH) & M),
               (int)((x.w >> SH) & M), 0, 0, 0, 0};
}

// E8M0 scale of plane j on each operand: w_j * w_j * 2^(2 e_j) = 1.
template <int J>
constexpr int scale() { return J == 0 ? 128 : J == 1 ? 127 : 126; }

// Plane J of one 256-bit k stage for a wave's TA x TB tiles of 32 x 32:
// expand the fragments (5 VALU per 4 dwords over the four planes), TA * TB
// MFMAs.
template <int J, int TA, int TB>
__device__ __forceinline__ void mm(const uint4 (&ra)[TA], const uint4 (&rb)[TB], v16f (&acc)[TA][TB]) {
    v8i fa[TA], fb[TB];
#pragma unroll
    for (int t = 0; t < TA; ++t) fa[t] = plane<J>(ra[t]);
#pragma unroll
    for (int t = 0; t < TB; ++t) fb[t] = plane<J>(rb[t]);
#pragma unroll
    for (int tr = 0; tr < TA; ++tr)
#pragma unroll
        for (int tc = 0; tc < TB; ++tc)
            acc[tr][tc] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                fa[tr], fb[tc], acc[tr][tc], 4, 4, 0, scale<J>(), 0, scale<J>());
}

// Parity words of one wave's 64 x 64 block of C (2 x 2 MFMA tiles a[q][h]
// whose accumulators started at 2^23: a count c < 2^23 leaves 2^23 + c
// exactly, so bit 0 of the float is the parity).  Each lane packs its 64
// parities into one word with one v_alignbit per bit (register r of lane l:
// row 8(r/4) + 4(l/32) + r%4, column l%32 of its tile), the in-wave 64 x 64
// bit transpose gives lane j the 64 lanes' bit j, i.e. 32 columns of two
// rows, and lanes j and j^16 (the two column tiles) combine their halves
// into one row's word, which lane j stores: row row0 + its row, word
// col0 / 64 of C, bits past column n cleared, nothing past row m or word nw.
__device__ __forceinline__ void store_parity_block(const v16f &a00, const v16f &a01, const v16f &a10,
                                                   const v16f &a11, int row0, int col0,
                                                   uint64_t *__restrict__ C, size_t ldc, int m, int n) {
    const int lane = threadIdx.x & 63;
    const int nw = (n + 63) >> 6;
    uint32_t xh[2] = {0u, 0u};
#pragma unroll
    for (int e = 0; e < 16; ++e) xh[0] = __builtin_amdgcn_alignbit(__float_as_uint(a00[e]), xh[0], 1);
#pragma unroll
    for (int e = 0; e < 16; ++e) xh[0] = __builtin_amdgcn_alignbit(__float_as_uint(a01[e]), xh[0], 1);
#pragma unroll
    for (int e = 0; e < 16; ++e) xh[1] = __builtin_amdgcn_alignbit(__float_as_uint(a10[e]), xh[1], 1);
#pragma unroll
    for (int e = 0; e < 16; ++e) xh[1] = __builtin_amdgcn_alignbit(__float_as_uint(a11[e]), xh[1], 1);
    const uint64_t y = dev::transpose64(((uint64_t)xh[1] << 32) | xh[0], lane);
    const uint64_t yp = dev::shfl_xor64(y, 16);
    const int tri = lane >> 5, tci = (lane >> 4) & 1, r = lane & 15;
    uint64_t word = tci == 0 ? ((y & 0xffffffffull) | (yp << 32)) : ((yp >> 32) | (y & 0xffffffff00000000ull));
    const int row = row0 + tri * 32 + 8 * (r >> 2) + (r & 3) + 4 * tci;
    const int w = col0 >> 6;
    if (w == nw - 1 && (n & 63)) word &= (1ull << (n & 63)) - 1ull;
    if (row < m && w < nw) C[(size_t)row * ldc + w] = word;
}

// Stage-tiled packed layout (PK): rows padded to 256; block b of 256 rows,
// stage s (256 bits of k): 8 KiB at (b * nst + s) * 8 KiB, row r's 32 bytes
// (k bits 256 s .. 256 s + 255, little-endian words) at r * 32.
constexpr int ROWS = 256;
constexpr int STAGE = 32;           // packed bytes of k per row per stage
constexpr int SEG = ROWS * STAGE;   // one stage of one 256-row block: 8 KiB

}  // namespace pk
}  // namespace mela
