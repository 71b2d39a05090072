// Bit packing, unpacking, bit-matrix transposition and the seeded generator.
//
// Packed layout (shared with the reference, matmul_gf2.c:33-36): row-major,
// `ld` 64-bit words per row, bit j of a row at word j>>6, bit j&63.
#include "common.h"
#include "transpose_device.h"

namespace mela {
namespace {

// ((int)x) & 1 with the x86-64 result for NaN/Inf/out-of-range (the
// reference's cast, e.g. matmul_gf2.c:66): cvttsd2si returns 0x80000000 there,
// whose low bit is 0.  v_cvt_i32_f64 would saturate instead (INT_MAX is odd),
// so the range test is explicit.
__device__ __forceinline__ uint32_t parity_of(double x) {
    if (!(x > -2147483649.0 && x < 2147483648.0)) return 0u;
    return (uint32_t)((int)x) & 1u;
}

constexpr int PACK_WPT = 4;  // packed words per thread

// One thread per (row, 4 consecutive words).  Reads of a column are
// coalesced across the rows of a wave; each thread writes 32 contiguous bytes.
template <typename T>
__global__ void __launch_bounds__(256) k_pack(const T *__restrict__ src, size_t rows,
                                              size_t cols, uint64_t *__restrict__ dst,
                                              size_t ld) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t w0 = (size_t)blockIdx.y * PACK_WPT;
    if (i >= rows) return;
#pragma unroll
    for (int q = 0; q < PACK_WPT; ++q) {
        size_t w = w0 + q;
        if (w >= ld) break;
        uint64_t word = 0;
        size_t j0 = w * 64;
        if (j0 < cols) {
            size_t jn = cols - j0 < 64 ? cols - j0 : 64;
            const T *col = src + j0 * rows + i;
            for (size_t b = 0; b < jn; ++b) {
                uint64_t bit;
                if constexpr (sizeof(T) == 8)
                    bit = parity_of(col[b * rows]);
                else
                    bit = col[b * rows] != 0;
                word |= bit << b;
            }
        }
        dst[i * ld + w] = word;
    }
}

// Compressed sparse column input (MATLAB's jc / ir / pr arrays): one thread
// per stored entry finds its column by binary search over jc (jc[0] = 0,
// non-decreasing, jc[n] = nnz, checked by the host) and sets its bit with an
// atomic OR, since entries of different columns share packed words.  Value
// classes: none (every stored entry is 1), logical (nonzero), double (the
// dense rule, ((int)x) & 1).  Row indices >= m are flagged in *err and
// skipped.
__global__ void __launch_bounds__(256)
k_pack_csc(const uint64_t *__restrict__ jc, const uint64_t *__restrict__ ir,
           const void *__restrict__ vals, int vcls, size_t nnz, size_t m, size_t n,
           uint64_t *__restrict__ dst, size_t ld, unsigned *__restrict__ err) {
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < nnz;
         k += (size_t)gridDim.x * 256) {
        size_t lo = 0, hi = n;  // jc[lo] <= k < jc[hi]
        while (hi - lo > 1) {
            const size_t mid = (lo + hi) >> 1;
            if (jc[mid] <= k) lo = mid; else hi = mid;
        }
        const uint64_t row = ir[k];
        if (row >= m) {
            atomicOr(err, 1u);
            continue;
        }
        uint32_t bit = 1u;
        if (vcls == MELA_LOGICAL) bit = ((const uint8_t *)vals)[k] != 0;
        else if (vcls == MELA_DOUBLE) bit = parity_of(((const double *)vals)[k]);
        if (bit)
            atomicOr((unsigned long long *)&dst[row * ld + (lo >> 6)], 1ull << (lo & 63));
    }
}

// One thread per (row, word): 64 byte stores, each coalesced across the
// consecutive rows of a wave (column-major logical output).
__global__ void __launch_bounds__(256) k_unpack(const uint64_t *__restrict__ src, size_t ld,
                                                size_t rows, size_t cols,
                                                uint8_t *__restrict__ dst) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t w = blockIdx.y;
    if (i >= rows) return;
    uint64_t word = src[i * ld + w];
    size_t j0 = w * 64;
    size_t jn = cols - j0 < 64 ? cols - j0 : 64;
    uint8_t *out = dst + j0 * rows + i;
    for (size_t b = 0; b < jn; ++b) out[b * rows] = (uint8_t)((word >> b) & 1u);
}

// dst (cols x rows, packed along rows) = transpose of src (rows x cols).
// One wave per 64x64 bit block; blockIdx.x = row block, blockIdx.y = word.
__global__ void __launch_bounds__(64) k_transpose(const uint64_t *__restrict__ src, size_t lds,
                                                  size_t rows, size_t cols,
                                                  uint64_t *__restrict__ dst, size_t ldd) {
    int lane = threadIdx.x;
    size_t bi = blockIdx.x, bw = blockIdx.y;
    size_t r = bi * 64 + lane;
    uint64_t x = r < rows ? src[r * lds + bw] : 0;
    x = dev::transpose64(x, lane);
    size_t c = bw * 64 + lane;
    if (c < cols) dst[c * ldd + bi] = x;
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) k_gen(uint64_t k1, uint64_t k2, uint64_t *__restrict__ dst,
                                             size_t rows, size_t cols, size_t ld) {
    size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= rows * ld) return;
    size_t i = idx / ld, w = idx % ld;
    size_t wc = (cols + 63) / 64;
    uint64_t v = 0;
    if (w < wc) {
        v = splitmix64(splitmix64(k1 + (uint64_t)(i * wc + w)) ^ k2);
        if (w == wc - 1 && (cols & 63)) v &= (1ull << (cols & 63)) - 1;
    }
    dst[idx] = v;
}

__global__ void __launch_bounds__(256) k_zero(uint4 *__restrict__ p, size_t n16) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ src,
                                                uint4 *__restrict__ dst, size_t n16) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * 256;
    for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ void k_zero_tail(uint8_t *__restrict__ p, size_t n) {
    size_t i = threadIdx.x;
    if (i < n) p[i] = 0;
}

uint64_t host_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

mela_status launch_pack(hipStream_t s, const void *src, int cls, size_t rows, size_t cols,
                        uint64_t *dst, size_t ld) {
    if (rows == 0 || ld == 0) return MELA_OK;
    dim3 grid((unsigned)((rows + 255) / 256), (unsigned)((ld + PACK_WPT - 1) / PACK_WPT));
    if (cls == MELA_DOUBLE)
        hipLaunchKernelGGL(k_pack<double>, grid, dim3(256), 0, s, (const double *)src, rows,
                           cols, dst, ld);
    else
        hipLaunchKernelGGL(k_pack<uint8_t>, grid, dim3(256), 0, s, (const uint8_t *)src, rows,
                           cols, dst, ld);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

mela_status launch_pack_csc(hipStream_t s, const uint64_t *jc, const uint64_t *ir,
                            const void *vals, int vcls, size_t nnz, size_t m, size_t n,
                            uint64_t *dst, size_t ld, unsigned *err) {
    if (nnz == 0) return MELA_OK;
    size_t blocks = (nnz + 255) / 256;
    if (blocks > 65536) blocks = 65536;  // grid-stride beyond that
    hipLaunchKernelGGL(k_pack_csc, dim3((unsigned)blocks), dim3(256), 0, s, jc, ir, vals, vcls, nnz,
                       m, n, dst, ld, err);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

mela_status launch_unpack(hipStream_t s, const uint64_t *src, size_t ld, size_t rows,
                          size_t cols, uint8_t *dst) {
    if (rows == 0 || cols == 0) return MELA_OK;
    dim3 grid((unsigned)((rows + 255) / 256), (unsigned)words64(cols));
    hipLaunchKernelGGL(k_unpack, grid, dim3(256), 0, s, src, ld, rows, cols, dst);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

mela_status launch_transpose(hipStream_t s, const uint64_t *src, size_t lds, size_t rows,
                             size_t cols, uint64_t *dst, size_t ldd) {
    if (rows == 0 || cols == 0) return MELA_OK;
    // every destination word (c, bi) with c < cols, bi < ceil(rows/64) is written
    dim3 grid((unsigned)words64(rows), (unsigned)words64(cols));
    hipLaunchKernelGGL(k_transpose, grid, dim3(64), 0, s, src, lds, rows, cols, dst, ldd);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

mela_status launch_gen_bernoulli(hipStream_t s, uint64_t seed, uint64_t *dst, size_t rows,
                                 size_t cols, size_t ld) {
    size_t total = rows * ld;
    if (total == 0) return MELA_OK;
    // two splitmix64 rounds: one round on a counter leaves GF(2)-linear
    // dependences between bit columns
    hipLaunchKernelGGL(k_gen, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       host_splitmix64(seed), host_splitmix64(seed ^ 0xD1B54A32D192ED03ull), dst,
                       rows, cols, ld);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

// Plain device copy (16-byte vectors); used to calibrate HBM byte counters.
mela_status launch_copy(hipStream_t s, void *dst, const void *src, size_t bytes) {
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return set_error(MELA_ERR_INVALID, "copy needs 16-byte aligned sizes and pointers");
    const size_t n16 = bytes / 16;
    if (!n16) return MELA_OK;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4 *)src,
                       (uint4 *)dst, n16);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

// Zero fill by kernel (not hipMemsetAsync, so callers may capture it in a graph).
mela_status launch_fill_zero(hipStream_t s, void *dst, size_t bytes) {
    if (bytes == 0) return MELA_OK;
    size_t n16 = bytes / 16;
    if (n16) {
        size_t blocks = (n16 + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint4 *)dst, n16);
    }
    size_t tail = bytes - n16 * 16;
    if (tail)
        hipLaunchKernelGGL(k_zero_tail, dim3(1), dim3(16), 0, s, (uint8_t *)dst + n16 * 16,
                           tail);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

}  // namespace mela
