// In-wave 64 x 64 bit-block transpose, shared by the packed transpose
// (pack.hip) and the fused transpose + expansion of the matrix-core product
// (mfma.hip).
#pragma once

#include "common.h"

namespace mela {
namespace dev {

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int s) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, s, 64);
    hi = __shfl_xor(hi, s, 64);
    return ((uint64_t)hi << 32) | lo;
}

// In-wave 64x64 bit transpose: lane l holds row l; afterwards lane c holds
// column c (bit l = old row l, bit c).  Six butterfly levels swap the
// off-diagonal s x s sub-blocks.
__device__ __forceinline__ uint64_t transpose64(uint64_t x, int lane) {
    const uint64_t masks[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull,
                               0x00FF00FF00FF00FFull, 0x0F0F0F0F0F0F0F0Full,
                               0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int lv = 0; lv < 6; ++lv) {
        int s = 32 >> lv;
        uint64_t m = masks[lv];
        uint64_t y = shfl_xor64(x, s);
        if ((lane & s) == 0)
            x = (x & m) | ((y & m) << s);
        else
            x = (x & ~m) | ((y & ~m) >> s);
    }
    return x;
}

}  // namespace dev
}  // namespace mela
