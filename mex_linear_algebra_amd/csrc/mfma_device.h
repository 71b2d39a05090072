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

// Stage-tiled packed layout (PK): rows padded to 256; block b of 256 rows,
// stage s (256 bits of k): 8 KiB at (b * nst + s) * 8 KiB, row r's 32 bytes
// (k bits 256 s .. 256 s + 255, little-endian words) at r * 32.
constexpr int ROWS = 256;
constexpr int STAGE = 32;           // packed bytes of k per row per stage
constexpr int SEG = ROWS * STAGE;   // one stage of one 256-row block: 8 KiB

}  // namespace pk
}  // namespace mela
