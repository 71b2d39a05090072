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
//
// Expanded operands are padded with zero rows to a multiple of 256 and with
// zero bytes to a row pitch that is a whole number of k stages, so the inner
// loops carry no bounds checks.
#include <cstdlib>
#include <functional>
#include <type_traits>

#include "common.h"
#include "mfma_device.h"
#include "transpose_device.h"

namespace mela {
namespace {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int XROWS = 256;         // row padding of the expanded operands
constexpr int RING = 4;            // k stages in the LDS ring of k_mfma_ring
constexpr int RING_STAGE = 64;     // bytes of k per row per ring stage
constexpr int SMALL_STAGE = 128;   // bytes of k per row per stage of k_mfma_small
constexpr int KPAD = 256;          // row pitch multiple: lcm(RING * RING_STAGE, SMALL_STAGE)
static_assert(KPAD % (RING * RING_STAGE) == 0 && KPAD % SMALL_STAGE == 0, "pitch multiple");

// Byte offset of k byte b of row i in an expanded operand of row pitch
// `pitch`.  Row-major for k_mfma_small.  Stage-tiled for k_mfma_ring: each
// block of XROWS rows stores its RING_STAGE-byte k stages one after another,
// a stage being XROWS rows x RING_STAGE bytes contiguous, so each of that
// kernel's DMA instructions reads 1 KiB of consecutive bytes (16 rows of a
// row-major operand sit 16 rows x pitch apart, and that form fed the matrix
// cores at 0.68x the rate of this one at 32768^3).  Writers keep each 32- or
// 64-byte chunk inside one stage segment (b % 32 == 0).
__device__ __forceinline__ size_t xoff(size_t i, size_t b, size_t pitch, bool tiled) {
    if (!tiled) return i * pitch + b;
    return (i / XROWS) * (XROWS * pitch) + (b / RING_STAGE) * (XROWS * RING_STAGE) +
           (i % XROWS) * RING_STAGE + b % RING_STAGE;
}

// bits -> expanded bytes, one thread per 32-byte output chunk.  FP4: each
// source bit becomes one nibble (0x0 or 0x2), two per byte (element 2i in the
// low nibble), one source word per chunk; INT8: one byte (0 or 1), half a
// source word per chunk.  Rows >= rows and words >= kw are written as zeros.
template <bool FP4>
__global__ void __launch_bounds__(256)
k_expand(const uint64_t *__restrict__ src, size_t lds, size_t rows, size_t kw, size_t rows_pad,
         size_t pitch, uint8_t *__restrict__ dst, Gate gate) {
    if (gate_skip(gate, false)) return;
    const size_t chunks = pitch / 32;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t i = idx / chunks, c = idx % chunks;
    if (i >= rows_pad) return;
    const size_t w = FP4 ? c : c >> 1;
    uint32_t o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < rows && w < kw) {
        const uint64_t x = src[i * lds + w];
        if (FP4) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {  // 8 bits -> one dword
                const uint32_t b = (uint32_t)(x >> (8 * q)) & 0xffu;
                uint32_t d = 0;
#pragma unroll
                for (int t = 0; t < 8; ++t) d |= ((b >> t) & 1u) << (4 * t + 1);
                o[q] = d;
            }
        } else {
            const uint32_t h = (uint32_t)(x >> (32 * (c & 1)));
#pragma unroll
            for (int q = 0; q < 8; ++q) {  // 4 bits -> one dword
                const uint32_t b = (h >> (4 * q)) & 0xfu;
                o[q] = (b & 1u) | ((b & 2u) << 7) | ((b & 4u) << 14) | ((b & 8u) << 21);
            }
        }
    }
    uint4 *p = reinterpret_cast<uint4 *>(dst + i * pitch + c * 32);
    p[0] = make_uint4(o[0], o[1], o[2], o[3]);
    p[1] = make_uint4(o[4], o[5], o[6], o[7]);
}

// One 64-bit source word -> its 32 (FP4) or 64 (INT8) expanded bytes as
// dwords, in the byte order of k_expand.
template <bool FP4>
__device__ __forceinline__ void expand_word(uint64_t x, uint32_t *o) {
    if constexpr (FP4) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t b = (uint32_t)(x >> (8 * q)) & 0xffu;
            uint32_t d = 0;
#pragma unroll
            for (int t = 0; t < 8; ++t) d |= ((b >> t) & 1u) << (4 * t + 1);
            o[q] = d;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t b = (uint32_t)(x >> (4 * q)) & 0xfu;
            o[q] = (b & 1u) | ((b & 2u) << 7) | ((b & 4u) << 14) | ((b & 8u) << 21);
        }
    }
}

// bits -> stage-tiled expanded bytes (the layout of xoff with tiled = true).
// One thread per RING_STAGE-byte segment of one row (two source words for
// FP4, one for INT8).  Four consecutive lanes take four consecutive segments
// of one row, so a wave reads 16 rows x 4 segments of source with each row's
// part contiguous and writes four runs of 16 rows x 64 bytes, each 1 KiB of
// consecutive output.  grid.x walks segment groups (neighbouring workgroups
// read neighbouring parts of the same rows), grid.y 64-row groups.  Rows >=
// rows and words >= kw are written as zeros.
template <bool FP4>
__global__ void __launch_bounds__(256)
k_expand_tiled(const uint64_t *__restrict__ src, size_t lds, size_t rows, size_t kw, size_t pitch,
               uint8_t *__restrict__ dst, Gate gate) {
    if (gate_skip(gate, false)) return;
    static_assert(XROWS % 64 == 0 && RING_STAGE == 64, "64-row groups inside a row block");
    const int l = threadIdx.x;
    const size_t seg = (size_t)blockIdx.x * 4 + (l & 3);
    const size_t i = (size_t)blockIdx.y * 64 + (l >> 2);
    constexpr int WPS = FP4 ? 2 : 1;  // source words per segment
    uint32_t o[16];
#pragma unroll
    for (int h = 0; h < WPS; ++h) {
        const size_t w = seg * WPS + h;
        const uint64_t x = (i < rows && w < kw) ? src[i * lds + w] : 0ull;
        expand_word<FP4>(x, o + 8 * h);
    }
    uint4 *p = reinterpret_cast<uint4 *>(dst + (i / XROWS) * (XROWS * pitch) +
                                         seg * (XROWS * RING_STAGE) + (i % XROWS) * RING_STAGE);
#pragma unroll
    for (int q = 0; q < 4; ++q) p[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
}

// B^T expanded straight from B: each wave transposes KB 64 x 64 bit blocks
// of B (64 k-rows each, 64 columns).  After the in-wave transpose lane c
// holds column 64*jb + c for those k values and writes their expansion to row
// 64*jb + c of the expanded B^T; KB = 2 for stage-tiled FP4, so that each
// lane writes one whole 64-byte stage segment.  Columns >= n and k rows >= k
// come out as zeros, which is the padding the product kernels expect.
template <bool FP4, int KB>
__global__ void __launch_bounds__(256)
k_transpose_expand(const uint64_t *__restrict__ B, size_t ldb, size_t k, size_t n, size_t kwp,
                   uint8_t *__restrict__ dst, size_t pitch, bool tiled, Gate gate) {
    if (gate_skip(gate, false)) return;
    static_assert(KB == 1 || (FP4 && KB == 2), "a lane writes at most 64 bytes");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t kb0 = ((size_t)blockIdx.x * 4 + wave) * KB;  // first 64-row block of B (k)
    const size_t jb = blockIdx.y;                               // 64-column block of B
    if (kb0 >= kwp) return;
    const size_t nw = (n + 63) >> 6;
    uint32_t o[16];
#pragma unroll
    for (int h = 0; h < KB; ++h) {
        const size_t i = (kb0 + h) * 64 + lane;
        uint64_t v = 0;
        if (i < k && jb < nw) {
            v = B[i * ldb + jb];
            if (jb == nw - 1 && (n & 63)) v &= (1ull << (n & 63)) - 1ull;
        }
        v = dev::transpose64(v, lane);
        expand_word<FP4>(v, o + 8 * h);
    }
    constexpr int NQ = FP4 ? 2 * KB : 4;  // 16-byte stores per lane
    uint4 *out = reinterpret_cast<uint4 *>(dst + xoff(jb * 64 + lane, kb0 * (FP4 ? 32 : 64), pitch, tiled));
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
}

// Workgroup -> C tile.  Workgroups are dealt to the 8 XCDs round-robin, so
// the first step gives each XCD a contiguous range of tile indices; the
// second walks those in groups of GR tile rows, column by column, so the
// workgroups resident on one XCD share A and B^T strips in its L2.  For
// k_mfma_ring (32 workgroups per XCD) GR = 4 keeps 4 x 8 tiles in flight per
// XCD: at 32768^3 its loads alone took 8.1 ms against 9.0 ms with GR = 8
// (the A strips are private to an XCD and are re-read half as often, the
// B^T strips are read by every XCD at about the same time).
template <int GR>
__device__ __forceinline__ void tile_of(int L, int tx, int ty, int &bx, int &by) {
    const int G = tx * ty;
    if ((G & 7) == 0) L = (L & 7) * (G >> 3) + (L >> 3);
    const int grp = L / (GR * ty);
    const int fx = grp * GR;
    const int gs = min(tx - fx, GR);
    const int r = L - grp * GR * ty;
    bx = fx + r % gs;
    by = r / gs;
}

template <bool FP4>
__device__ __forceinline__ void mfma_step(const uint4 &fa, const uint4 &fb,
                                          typename std::conditional<FP4, v16f, v16i>::type &acc) {
    if constexpr (FP4) {
        v8i a8 = {(int)fa.x, (int)fa.y, (int)fa.z, (int)fa.w, 0, 0, 0, 0};
        v8i b8 = {(int)fb.x, (int)fb.y, (int)fb.z, (int)fb.w, 0, 0, 0, 0};
        acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8, b8, acc, 4, 4, 0, 127, 0, 127);
    } else {
        v4i a4 = {(int)fa.x, (int)fa.y, (int)fa.z, (int)fa.w};
        v4i b4 = {(int)fb.x, (int)fb.y, (int)fb.z, (int)fb.w};
        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a4, b4, acc, 0, 0, 0);
    }
}

// C tile epilogue.  32x32 accumulator layout: register r of lane l holds row
// 8*(r/4) + 4*(l/32) + r%4 and column l%32 of its tile.  Column tiles
// (2c, 2c+1) of a wave form one 64-bit word per row; bits past column n of
// the last word are cleared.
template <int TR, typename Acc>
__device__ __forceinline__ void store_tile(Acc (&acc)[TR][TR], int row0, int col0,
                                           uint64_t *__restrict__ C, size_t ldc, int m, int n) {
    const int lane = threadIdx.x & 63;
    const int nw = (n + 63) >> 6;
    const uint64_t lastmask = (n & 63) ? ((1ull << (n & 63)) - 1ull) : ~0ull;
#pragma unroll
    for (int tr = 0; tr < TR; ++tr)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int rbase = row0 + tr * 32 + 8 * (r >> 2) + (r & 3);
#pragma unroll
            for (int tc = 0; tc < TR; tc += 2) {
                const uint64_t b0 = __ballot((int)acc[tr][tc][r] & 1);
                const uint64_t b1 = __ballot((int)acc[tr][tc + 1][r] & 1);
                const int w = (col0 + tc * 32) >> 6;
                // lane 0 writes the row of lanes 0-31, lane 32 that of lanes 32-63
                if ((lane & 31) == 0 && w < nw) {
                    const int sh = lane >> 5;
                    const int row = rbase + 4 * sh;
                    const uint64_t lo = sh ? (b0 >> 32) : (b0 & 0xffffffffull);
                    const uint64_t hi = sh ? (b1 >> 32) : (b1 & 0xffffffffull);
                    uint64_t word = lo | (hi << 32);
                    if (w == nw - 1) word &= lastmask;
                    if (row < m) C[(size_t)row * ldc + w] = word;
                }
            }
        }
}

// C tile epilogue of the direct kernel, whose accumulators start at 2^23:
// the wave's 128 x 128 as four 64 x 64 parity blocks (pk::store_parity_block).
template <int TR>
__device__ __forceinline__ void store_tile_offset(v16f (&acc)[TR][TR], int row0, int col0,
                                                  uint64_t *__restrict__ C, size_t ldc, int m, int n) {
    static_assert(TR % 2 == 0, "quarters of 2 x 2 MFMA tiles");
#pragma unroll
    for (int cr = 0; cr < TR / 2; ++cr)
#pragma unroll
        for (int cc = 0; cc < TR / 2; ++cc)
            pk::store_parity_block(acc[2 * cr][2 * cc], acc[2 * cr][2 * cc + 1], acc[2 * cr + 1][2 * cc],
                                   acc[2 * cr + 1][2 * cc + 1], row0 + cr * 64, col0 + cc * 64, C, ldc, m, n);
}

// ---------------------------------------------------------------------------
// k_mfma_small: 128 x 128 tiles for grids too small to fill the chip with
// 256 x 256 tiles.  4 waves in a 2 x 2 grid, each 64 x 64 (2 x 2 MFMA tiles);
// two workgroups per CU.  Stage = 128 bytes of k, staged through registers:
// thread t loads 16-byte unit t%8 of rows t/8 + 32q, so eight lanes read one
// 128-byte row segment.  LDS is fragment-major with the row index
// XOR-swizzled by f(u) = u ^ ((u&1)<<3) (< 16, so every 32-row block is
// permuted within itself and fragment reads stay contiguous, while the eight
// units of a row land in distinct banks).
template <bool FP4>
__global__ void __launch_bounds__(256, 2)
k_mfma_small(const uint8_t *__restrict__ A, const uint8_t *__restrict__ B, size_t pitch,
             uint64_t *__restrict__ C, size_t ldc, int m, int n, int tx, int ty, Gate gate) {
    if (gate_skip(gate, false)) return;
    constexpr int TR = 2, WT = 64, ROWS = 128, U = 8;
    __shared__ uint4 sA[2][U * ROWS];
    __shared__ uint4 sB[2][U * ROWS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    int bx, by;
    tile_of<8>(blockIdx.x, tx, ty, bx, by);
    const int i0 = bx * ROWS, j0 = by * ROWS;
    const int lu = tid & 7, lr = tid >> 3;
    const int fsw = lu ^ ((lu & 1) << 3);
    const uint8_t *pa = A + (size_t)(i0 + lr) * pitch + lu * 16;
    const uint8_t *pb = B + (size_t)(j0 + lr) * pitch + lu * 16;
    const size_t rs = 32 * pitch;
    uint4 xa0, xa1, xa2, xa3, xb0, xb1, xb2, xb3;
    auto load = [&](int st) __attribute__((always_inline)) {
        const size_t ko = (size_t)st * SMALL_STAGE;
        xa0 = *reinterpret_cast<const uint4 *>(pa + ko);
        xb0 = *reinterpret_cast<const uint4 *>(pb + ko);
        xa1 = *reinterpret_cast<const uint4 *>(pa + rs + ko);
        xb1 = *reinterpret_cast<const uint4 *>(pb + rs + ko);
        xa2 = *reinterpret_cast<const uint4 *>(pa + 2 * rs + ko);
        xb2 = *reinterpret_cast<const uint4 *>(pb + 2 * rs + ko);
        xa3 = *reinterpret_cast<const uint4 *>(pa + 3 * rs + ko);
        xb3 = *reinterpret_cast<const uint4 *>(pb + 3 * rs + ko);
    };
    auto store = [&](int buf) __attribute__((always_inline)) {
        const int s0 = lu * ROWS + (lr ^ fsw);
        sA[buf][s0] = xa0;
        sB[buf][s0] = xb0;
        sA[buf][s0 + 32] = xa1;
        sB[buf][s0 + 32] = xb1;
        sA[buf][s0 + 64] = xa2;
        sB[buf][s0 + 64] = xb2;
        sA[buf][s0 + 96] = xa3;
        sB[buf][s0 + 96] = xb3;
    };
    using Acc = typename std::conditional<FP4, v16f, v16i>::type;
    Acc acc[TR][TR];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;
    const int half = lane >> 5, l32 = lane & 31;
    auto rd = [&](int buf, int kb, uint4 *fa, uint4 *fb) __attribute__((always_inline)) {
        const int u = 2 * kb + half;
        const int fu = u ^ (half << 3);
        const int ra = u * ROWS + wr * WT + (l32 ^ fu);
        const int rb = u * ROWS + wc * WT + (l32 ^ fu);
#pragma unroll
        for (int t = 0; t < TR; ++t) {
            fa[t] = sA[buf][ra + t * 32];
            fb[t] = sB[buf][rb + t * 32];
        }
    };
    auto mm = [&](const uint4 *fa, const uint4 *fb) __attribute__((always_inline)) {
#pragma unroll
        for (int tr = 0; tr < TR; ++tr)
#pragma unroll
            for (int tc = 0; tc < TR; ++tc) mfma_step<FP4>(fa[tr], fb[tc], acc[tr][tc]);
    };
    // fragments of k-block kb+1 are read before the MFMAs of k-block kb
    auto compute = [&](int buf) __attribute__((always_inline)) {
        uint4 pa0[TR], pb0[TR], pa1[TR], pb1[TR];
        rd(buf, 0, pa0, pb0);
        __builtin_amdgcn_sched_barrier(0);
        rd(buf, 1, pa1, pb1);
        __builtin_amdgcn_sched_barrier(0);
        mm(pa0, pb0);
        __builtin_amdgcn_sched_barrier(0);
        rd(buf, 2, pa0, pb0);
        __builtin_amdgcn_sched_barrier(0);
        mm(pa1, pb1);
        __builtin_amdgcn_sched_barrier(0);
        rd(buf, 3, pa1, pb1);
        __builtin_amdgcn_sched_barrier(0);
        mm(pa0, pb0);
        __builtin_amdgcn_sched_barrier(0);
        mm(pa1, pb1);
    };
    const int nstage = (int)(pitch / SMALL_STAGE);
    load(0);
    store(0);
    __syncthreads();
    for (int st = 0; st < nstage; ++st) {
        const int buf = st & 1;
        // the last prefetch re-reads a valid stage into the idle buffer, so
        // the loop body stays one basic block and the barriers keep its order
        load(st + 1 < nstage ? st + 1 : st);
        __builtin_amdgcn_sched_barrier(0);
        compute(buf);
        __builtin_amdgcn_sched_barrier(0);
        store(buf ^ 1);
        __syncthreads();
    }
    store_tile<TR>(acc, i0 + wr * WT, j0 + wc * WT, C, ldc, m, n);
}

// ---------------------------------------------------------------------------
// k_mfma_ring: 256 x 256 tiles, 4 waves in a 2 x 2 grid, each 128 x 128
// (4 x 4 MFMA tiles, 256 accumulator registers), one workgroup per CU.
//
// Operands are stage-tiled (xoff): the 256 rows x RING_STAGE bytes of one
// stage of one operand are 16 KiB of consecutive bytes.  They move from HBM
// straight into a RING-stage LDS ring with global_load_lds_dwordx4 (no
// staging registers, 1 KiB contiguous per instruction), so RING-2 stages are in
// flight while the MFMAs work on the current one.  A stage holds RING_STAGE
// bytes of k for the 256 A rows and the 256 B^T rows; each wave issues
// exactly 2U DMA instructions per stage, so "stage s landed" is a plain
// vmcnt threshold.  The ring buffers are separate LDS arrays and the loop is
// unrolled by RING, so every LDS access names its buffer statically and the
// compiler does not make fragment reads wait for the ring's other stages.
//
// Stage layout: row-major, U = RING_STAGE/16 units per row; unit p of row r
// holds global unit p ^ g(r) with g(r) = (r>>2) & 3, so four lanes still
// read one 64-byte row segment.  Measured on gfx950 (ds_read_b128, four
// waves per CU): fragment reads with this g run at the rate of 64 contiguous
// 16-byte slots (88 B/clk/CU), while g(r) = r & 3, no swizzle and
// ((r>>1) ^ (r>>3)) & 3 all drop to 64 B/clk.
//
// Step s: wait until stage s+1 landed, barrier, refill the buffer of stage
// s-1 with stage s+RING-1, then read k-block 1 of stage s and k-block 0 of
// stage s+1 from LDS into registers while the MFMAs consume the fragments
// read before.  Fragment reads are interleaved with the first MFMAs of each
// block, the refill's DMA instructions with both blocks.
__device__ __forceinline__ int ring_swz(int r) { return (r >> 2) & 3; }

template <bool FP4>
__global__ void __launch_bounds__(256, 1)
k_mfma_ring(const uint8_t *__restrict__ A, const uint8_t *__restrict__ B, size_t pitch,
            uint64_t *__restrict__ C, size_t ldc, int m, int n, int tx, int ty, Gate gate) {
    if (gate_skip(gate, false)) return;
    // The host pass only needs the launch stub; the body uses LDS-DMA
    // builtins that exist for the device target alone.
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr int TR = 4, WT = 128, ROWS = 256, U = RING_STAGE / 16;
    constexpr int SS = ROWS * U;  // uint4 slots per operand per stage
    static_assert(RING == 4 && U == 4, "ring layout");
    __shared__ uint4 r0[2 * SS], r1[2 * SS], r2[2 * SS], r3[2 * SS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    int bx, by;
    tile_of<4>(blockIdx.x, tx, ty, bx, by);
    const int i0 = bx * ROWS, j0 = by * ROWS;
    // DMA sources: instruction i (0..U-1) of this wave fills slots
    // (i*4 + wave)*64 + lane of a stage (row S/U, position S%U)
    const uint8_t *da[U], *db[U];
#pragma unroll
    for (int i = 0; i < U; ++i) {
        const int S = (i * 4 + wave) * 64 + lane;
        const int r = S / U, p = S % U;
        const int u = p ^ ring_swz(r);
        da[i] = A + (size_t)i0 * pitch + r * RING_STAGE + u * 16;
        db[i] = B + (size_t)j0 * pitch + r * RING_STAGE + u * 16;
    }
    const int nstage = (int)(pitch / RING_STAGE);  // a multiple of RING
#define MELA_LDS(x) ((__attribute__((address_space(3))) void *)(x))
    // DMA instructions i0 .. i1-1 (of U) of stage st for both operands
    auto dma = [&](uint4 *buf, int st, int i0, int i1) __attribute__((always_inline)) {
        const size_t ko = (size_t)(st < nstage ? st : nstage - 1) * (ROWS * RING_STAGE);
#pragma unroll
        for (int i = i0; i < i1; ++i) {
            __builtin_amdgcn_global_load_lds(da[i] + ko, MELA_LDS(buf + (i * 4 + wave) * 64), 16, 0, 0);
            __builtin_amdgcn_global_load_lds(db[i] + ko, MELA_LDS(buf + SS + (i * 4 + wave) * 64), 16, 0, 0);
        }
    };
#undef MELA_LDS
    using Acc = typename std::conditional<FP4, v16f, v16i>::type;
    Acc acc[TR][TR];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;
    const int half = lane >> 5, l32 = lane & 31;
    auto rdA = [&](const uint4 *buf, int kb, uint4 *fa) __attribute__((always_inline)) {
        const int u = 2 * kb + half;
#pragma unroll
        for (int t = 0; t < TR; ++t) {
            const int ra = wr * WT + t * 32 + l32;
            fa[t] = buf[ra * U + (u ^ ring_swz(ra))];
        }
    };
    auto rdB = [&](const uint4 *buf, int kb, uint4 *fb) __attribute__((always_inline)) {
        const int u = 2 * kb + half;
#pragma unroll
        for (int t = 0; t < TR; ++t) {
            const int rb = wc * WT + t * 32 + l32;
            fb[t] = buf[SS + rb * U + (u ^ ring_swz(rb))];
        }
    };
    auto mmh = [&](const uint4 *fa, const uint4 *fb, int t0) __attribute__((always_inline)) {
#pragma unroll
        for (int tr = t0; tr < t0 + 2; ++tr)
#pragma unroll
            for (int tc = 0; tc < TR; ++tc) mfma_step<FP4>(fa[tr], fb[tc], acc[tr][tc]);
    };
    // s_waitcnt immediates (gfx9 encoding): vmcnt(V) and lgkmcnt(L), expcnt free
    constexpr int VMC = 2 * U * (RING - 3);  // in flight once stage s+1 landed
    constexpr int WAIT = (VMC & 0xF) | ((VMC >> 4) << 14) | (7 << 4) | (0 << 8);
    constexpr int V0 = 2 * U * (RING - 2);   // in flight once stage 0 landed
    constexpr int WAIT0 = (V0 & 0xF) | ((V0 >> 4) << 14) | (7 << 4) | (0xF << 8);
    uint4 fa0[TR], fb0[TR], fa1[TR], fb1[TR];
    // step st: fa0/fb0 hold k-block 0 of stage st
    // The refill's 2U DMA instructions are spread over both halves of the
    // step, one after every second MFMA of the interleave, rather than issued
    // in one burst: a wave stalls at issue once the address unit's queue is
    // full, and its MFMAs wait behind it.
    auto step = [&](const uint4 *cur, const uint4 *nx1, uint4 *refill, int st) __attribute__((always_inline)) {
        __builtin_amdgcn_s_waitcnt(WAIT);
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        dma(refill, st + RING - 1, 0, U / 2);
        rdA(cur, 1, fa1);
        rdB(cur, 1, fb1);
        mmh(fa0, fb0, 0);
        mmh(fa0, fb0, 2);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                 // MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                 // LDS read
            if (i % 2 == 1) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0); // VMEM read (DMA)
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        dma(refill, st + RING - 1, U / 2, U);
        rdA(nx1, 0, fa0);
        rdB(nx1, 0, fb0);
        mmh(fa1, fb1, 0);
        mmh(fa1, fb1, 2);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 1);
            if (i % 2 == 1) __builtin_amdgcn_sched_group_barrier(0x020, 1, 1);
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    dma(r0, 0, 0, U);
    dma(r1, 1, 0, U);
    dma(r2, 2, 0, U);
    __builtin_amdgcn_s_waitcnt(WAIT0);
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_sched_barrier(0);
    rdA(r0, 0, fa0);
    rdB(r0, 0, fb0);
    __builtin_amdgcn_sched_barrier(0);
    // stage j lives in buffer j % RING; DMAs past the last stage re-read it
    // into buffers that are never computed on
    for (int st = 0; st < nstage; st += RING) {
        step(r0, r1, r3, st);
        step(r1, r2, r0, st + 1);
        step(r2, r3, r1, st + 2);
        step(r3, r0, r2, st + 3);
    }
    // the trailing DMAs must land before the workgroup's LDS is released
    __builtin_amdgcn_s_waitcnt(0);
    store_tile<TR>(acc, i0 + wr * WT, j0 + wc * WT, C, ldc, m, n);
#endif
}

// ---------------------------------------------------------------------------
// Packed operands, expanded in registers.
//
// The kernels above stream operands that were expanded to 4 bits per k in
// HBM; here the matrix cores are fed from packed bits (1 bit per k), so HBM,
// the L2 and the L1 carry 4x fewer bytes and no expanded copy exists.  The
// expansion is nearly free because of how e2m1 encodes small values
// (mfma_device.h: four masked planes per packed dword, unit products by the
// E8M0 block scales, C(i,j) = acc & 1).
//
// Stage-tiled packed layout (PK, mfma_device.h): rows padded to 256, k padded
// to a whole number of loop trips (PK_DIRECT stages of 256 bits); each
// wave-wide fragment load reads 1 KiB of consecutive bytes.
constexpr int PK_STAGE = 32;                 // packed bytes of k per row per stage (256 k)
constexpr int PK_SEG = XROWS * PK_STAGE;     // one stage of one 256-row block: 8 KiB
#ifndef MELA_PK_DIRECT
#define MELA_PK_DIRECT 4
#endif
constexpr int PK_DIRECT = MELA_PK_DIRECT;    // stages in flight in k_mfma_direct

// A (rows x k bits, lds words per row) -> PK layout.  One thread per row of
// a 256-row block and TP_STAGES consecutive stages: it reads 256 contiguous
// source bytes (16-byte loads when `vec`: src 16-byte aligned, lds even) and
// writes its 32 bytes of each stage segment, so a wave's stores are runs of
// 64 rows x 32 bytes.  (One thread per (row, stage), four lanes per row, ran
// at 2.5 TB/s against 5.9 TB/s for this form at 32768 x 32768.)  grid.x:
// groups of TP_STAGES stages (the last one partial when nst is not a
// multiple), grid.y: 256-row blocks.  Rows >= rows and words >= kw are
// written as zeros.  count != nullptr: also adds the number of set bits of A
// to the GATE_SLOTS counts at count (the density gate of the AUTO product;
// one atomic per workgroup).
constexpr int TP_STAGES = 8;
__global__ void __launch_bounds__(256)
k_tile_packed(const uint64_t *__restrict__ src, size_t lds, size_t rows, size_t kw, size_t nst, bool vec,
              uint8_t *__restrict__ dst, unsigned long long *__restrict__ count) {
    const int l = threadIdx.x;
    const size_t i = (size_t)blockIdx.y * XROWS + l;
    const size_t s0 = (size_t)blockIdx.x * TP_STAGES;
    const uint64_t *row = src + i * lds + 4 * s0;
    uint4 v[2 * TP_STAGES];
    if (vec && i < rows && 4 * (s0 + TP_STAGES) <= kw) {
#pragma unroll
        for (int t = 0; t < 2 * TP_STAGES; ++t) v[t] = reinterpret_cast<const uint4 *>(row)[t];
    } else {
#pragma unroll
        for (int t = 0; t < 2 * TP_STAGES; ++t) {
            const size_t w = 4 * s0 + 2 * t;
            const uint64_t x = (i < rows && w < kw) ? row[2 * t] : 0ull;
            const uint64_t y = (i < rows && w + 1 < kw) ? row[2 * t + 1] : 0ull;
            v[t] = make_uint4((uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32));
        }
    }
    unsigned long long c = 0;
#pragma unroll
    for (int q = 0; q < TP_STAGES; ++q) {
        if (s0 + q < nst) {
            uint4 *p = reinterpret_cast<uint4 *>(dst + (blockIdx.y * nst + s0 + q) * PK_SEG + (size_t)l * PK_STAGE);
            p[0] = v[2 * q];
            p[1] = v[2 * q + 1];
        }
        if (count)
            c += (unsigned long long)(__popc(v[2 * q].x) + __popc(v[2 * q].y) + __popc(v[2 * q].z) +
                                      __popc(v[2 * q].w) + __popc(v[2 * q + 1].x) + __popc(v[2 * q + 1].y) +
                                      __popc(v[2 * q + 1].z) + __popc(v[2 * q + 1].w));
    }
    if (count) {
        __shared__ unsigned long long part[4];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if ((l & 63) == 0) part[l >> 6] = c;
        __syncthreads();
        if (l == 0)
            atomicAdd(count + (blockIdx.y * gridDim.x + blockIdx.x) % GATE_SLOTS,
                      part[0] + part[1] + part[2] + part[3]);
    }
}

// B (k x n bits) -> B^T in PK layout.  Each wave transposes the 64 x 64 bit
// blocks of one stage (256 k rows) for four column words: a lane reads 32
// contiguous bytes of each of its four k rows (16-byte loads when `vec`: B
// 16-byte aligned, ldb even), and after the in-wave transposes lane c holds
// column 64 jb + c's 256 k bits of each word jb, which it writes as its
// row's 32 bytes (a wave writes four runs of 2 KiB).  (One column word per
// wave ran at 3.1 TB/s against 5.4 TB/s for this form at 32768 x 32768.)
// grid.x: groups of 4 stages (one per wave), grid.y: 256-column blocks of
// the padded width.  Columns >= n and k rows >= k come out as zeros.
__global__ void __launch_bounds__(256)
k_transpose_tile_packed(const uint64_t *__restrict__ B, size_t ldb, size_t k, size_t n, size_t nst, bool vec,
                        uint8_t *__restrict__ dst, Gate gate) {
    if (gate_skip(gate, false)) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t s = (size_t)blockIdx.x * 4 + wave;
    const size_t jb = 4 * (size_t)blockIdx.y;
    if (s >= nst) return;
    const size_t nw = (n + 63) >> 6;
    uint64_t v[4][4];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const size_t i = (4 * s + h) * 64 + lane;
        const uint64_t *row = B + i * ldb + jb;
        if (vec && i < k && jb + 3 < nw) {
            const uint4 q0 = reinterpret_cast<const uint4 *>(row)[0];
            const uint4 q1 = reinterpret_cast<const uint4 *>(row)[1];
            v[0][h] = ((uint64_t)q0.y << 32) | q0.x;
            v[1][h] = ((uint64_t)q0.w << 32) | q0.z;
            v[2][h] = ((uint64_t)q1.y << 32) | q1.x;
            v[3][h] = ((uint64_t)q1.w << 32) | q1.z;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c][h] = (i < k && jb + c < nw) ? row[c] : 0ull;
        }
    }
    if (n & 63) {
        const uint64_t lastmask = (1ull << (n & 63)) - 1ull;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (jb + c == nw - 1)
#pragma unroll
                for (int h = 0; h < 4; ++h) v[c][h] &= lastmask;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int h = 0; h < 4; ++h) v[c][h] = dev::transpose64(v[c][h], lane);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const size_t j = (jb + c) * 64 + lane;
        uint4 *p = reinterpret_cast<uint4 *>(dst + ((j / XROWS) * nst + s) * PK_SEG + (j % XROWS) * PK_STAGE);
        p[0] = make_uint4((uint32_t)v[c][0], (uint32_t)(v[c][0] >> 32), (uint32_t)v[c][1], (uint32_t)(v[c][1] >> 32));
        p[1] = make_uint4((uint32_t)v[c][2], (uint32_t)(v[c][2] >> 32), (uint32_t)v[c][3], (uint32_t)(v[c][3] >> 32));
    }
}

// Plane J of a wave's fragments (mfma_device.h's plane map).
template <int J, int T>
__device__ __forceinline__ void planes(const uint4 (&r)[T], pk::v8i (&o)[T]) {
#pragma unroll
    for (int t = 0; t < T; ++t) o[t] = pk::plane<J>(r[t]);
}

// The TR x TR MFMAs of plane J from extracted operands.
template <int J, int TA, int TB>
__device__ __forceinline__ void mfmas(const pk::v8i (&fa)[TA], const pk::v8i (&fb)[TB], pk::v16f (&acc)[TA][TB]) {
#pragma unroll
    for (int tr = 0; tr < TA; ++tr)
#pragma unroll
        for (int tc = 0; tc < TB; ++tc)
            acc[tr][tc] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                fa[tr], fb[tc], acc[tr][tc], 4, 4, 0, pk::scale<J>(), 0, pk::scale<J>());
}

// Scheduling pattern of a region: NM times one MFMA, then NV vector ALU
// instructions.
template <int NM, int NV>
__device__ __forceinline__ void interleave() {
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);
    }
}

// k_mfma_direct: 256 x 256 tiles, 4 waves in a 2 x 2 grid, each 128 x 128
// (4 x 4 MFMA tiles, 256 accumulators), one workgroup per CU, the fragments
// loaded straight into registers.  In the PK layout lane l's
// fragment of row tile t (row r = 128 w + 32 t + l % 32, k half l / 32) is 16
// consecutive bytes at r * 32 + (l / 32) * 16 of the stage segment, so one
// global_load_dwordx4 per fragment moves 1 KiB of consecutive bytes per wave;
// the two waves that share a row block meet in the L1.  No LDS, no barrier:
// every wave waits only for its own loads, D stages ahead.  (An LDS ring
// filled by global_load_lds, one barrier per stage, ran 9.9 ms at 32768^3
// against 9.0 ms for this form.)  The accumulators start at 2^23, so the
// epilogue (store_tile_offset) reads each parity from bit 0 of its float;
// the host sends k >= 2^23 to the expanded-operand kernels.
template <int D>
__global__ void __launch_bounds__(256, 1)
k_mfma_direct(const uint8_t *__restrict__ A, const uint8_t *__restrict__ B, int nst,
              uint64_t *__restrict__ C, size_t ldc, int m, int n, int tx, int ty, Gate gate) {
    if (gate_skip(gate, false)) return;
    constexpr int TR = 4, WT = 128;
    constexpr int SEG16 = PK_SEG / 16;  // uint4 per 256-k segment of one 256-row block
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    int bx, by;
    tile_of<4>(blockIdx.x, tx, ty, bx, by);
    const int i0 = bx * XROWS, j0 = by * XROWS;
    const int half = lane >> 5, l32 = lane & 31;
    const uint4 *pa = reinterpret_cast<const uint4 *>(A + (size_t)bx * nst * PK_SEG) +
                      (wr * WT + l32) * 2 + half;
    const uint4 *pb = reinterpret_cast<const uint4 *>(B + (size_t)by * nst * PK_SEG) +
                      (wc * WT + l32) * 2 + half;
    v16f acc[TR][TR];
#pragma unroll
    for (int a = 0; a < TR; ++a)
#pragma unroll
        for (int b = 0; b < TR; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 8388608.0f;  // 2^23: parity in bit 0
    uint4 fa[D][TR], fb[D][TR];
    // stage st's fragments (past the last stage: the last one again)
    auto load = [&](int st, uint4 (&ra)[TR], uint4 (&rb)[TR]) __attribute__((always_inline)) {
        const size_t so = (size_t)(st < nst ? st : nst - 1) * SEG16;
#pragma unroll
        for (int t = 0; t < TR; ++t) {
            ra[t] = pa[so + t * 64];
            rb[t] = pb[so + t * 64];
        }
    };
#pragma unroll
    for (int j = 0; j < D; ++j) load(j, fa[j], fb[j]);
    __builtin_amdgcn_sched_barrier(0);
    // Each plane's 16 MFMAs run beside the extraction of the next plane's
    // operands (even planes in xa/xb, odd ones in qa/qb); the last plane's
    // run beside the refill of the stage's fragments and the extraction of
    // the next stage's plane 0, so the matrix core never waits for the
    // vector ALU at a stage boundary.  The host pads k to whole trips of D
    // stages; the group barriers pin every instruction class to its slot
    // (the scheduler would otherwise sink the loads to their use D steps
    // later and wait for them there).
    pk::v8i xa[TR], xb[TR], qa[TR], qb[TR];
    planes<0>(fa[0], xa);
    planes<0>(fb[0], xb);
    __builtin_amdgcn_sched_barrier(0);
    for (int st = 0; st < nst; st += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            planes<1>(fa[j], qa);
            planes<1>(fb[j], qb);
            mfmas<0>(xa, xb, acc);
            interleave<16, 2>();
            __builtin_amdgcn_sched_barrier(0);
            planes<2>(fa[j], xa);
            planes<2>(fb[j], xb);
            mfmas<1>(qa, qb, acc);
            interleave<16, 2>();
            __builtin_amdgcn_sched_barrier(0);
            planes<3>(fa[j], qa);
            planes<3>(fb[j], qb);
            mfmas<2>(xa, xb, acc);
            interleave<16, 4>();
            __builtin_amdgcn_sched_barrier(0);
            load(st + j + D, fa[j], fb[j]);
            planes<0>(fa[(j + 1) % D], xa);
            planes<0>(fb[(j + 1) % D], xb);
            mfmas<3>(qa, qb, acc);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
                __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);  // VALU
            }
            interleave<8, 4>();
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    store_tile_offset<TR>(acc, i0 + wr * WT, j0 + wc * WT, C, ldc, m, n);
}

}  // namespace

// C = A * B on the matrix cores from packed operands (fp4, >= 256 tiles of
// 256 x 256): A and B^T are re-laid in the stage-tiled packed layout (two
// passes over 2 x 128 MiB at 32768^3), then k_mfma_direct.
mela_status matmul_mfma_packed(mela_ctx *ctx, const uint64_t *A, size_t lda, const uint64_t *B,
                               size_t ldb, uint64_t *C, size_t ldc, size_t m, size_t k, size_t n,
                               Gate gate, unsigned long long *count,
                               const std::function<mela_status(Gate)> &sparse_path) {
    hipStream_t s = ctx->stream;
    const size_t kw = words64(k);
    const size_t q = PK_DIRECT;  // the kernel's loop runs whole trips of PK_DIRECT stages
    const size_t nst = ((k + 255) / 256 + q - 1) / q * q;
    const size_t mp = (m + XROWS - 1) / XROWS * XROWS, np = (n + XROWS - 1) / XROWS * XROWS;
    if (nst > (size_t)INT32_MAX || mp / 64 > 65535 || np / 64 > 65535)
        return set_error(MELA_ERR_INVALID, "matrix too large for the matrix-core product");
    void *pa = nullptr, *pb = nullptr;
    mela_status st;
    if ((st = scratch_get(ctx, SLOT_XA, mp * nst * PK_STAGE, &pa)) != MELA_OK) return st;
    if ((st = scratch_get(ctx, SLOT_XB, np * nst * PK_STAGE, &pb)) != MELA_OK) return st;
    const bool vec_a = ((uintptr_t)A & 15) == 0 && (lda & 1) == 0;
    const bool vec_b = ((uintptr_t)B & 15) == 0 && (ldb & 1) == 0;
    hipLaunchKernelGGL(k_tile_packed, dim3((unsigned)((nst + TP_STAGES - 1) / TP_STAGES), (unsigned)(mp / XROWS)),
                       dim3(256), 0, s, A, lda, m, kw, nst, vec_a, (uint8_t *)pa, count);
    if (count && sparse_path) {
        // the gate is complete once A is tiled: the row-XOR product goes here
        mela_status rs = sparse_path(gate);
        if (rs != MELA_OK) return rs;
    }
    hipLaunchKernelGGL(k_transpose_tile_packed, dim3((unsigned)((nst + 3) / 4), (unsigned)(np / XROWS)), dim3(256),
                       0, s, B, ldb, k, n, nst, vec_b, (uint8_t *)pb, gate);
    MELA_HIP(hipGetLastError());
    const size_t tx = mp / XROWS, ty = np / XROWS;
    if (tx * ty > (size_t)INT32_MAX) return set_error(MELA_ERR_INVALID, "too many tiles for one launch");
    prof_mark(ctx, true);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mfma_direct<PK_DIRECT>), dim3((unsigned)(tx * ty)), dim3(256), 0, s,
                       (const uint8_t *)pa, (const uint8_t *)pb, (int)nst, C, ldc, (int)m, (int)n, (int)tx,
                       (int)ty, gate);
    prof_mark(ctx, false);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

// C = A * B via the matrix cores; fp4 or int8 expansion.
mela_status matmul_mfma(mela_ctx *ctx, const uint64_t *A, size_t lda, const uint64_t *B,
                        size_t ldb, uint64_t *C, size_t ldc, size_t m, size_t k, size_t n,
                        bool fp4, Gate gate, unsigned long long *count,
                        const std::function<mela_status(Gate)> &sparse_path) {
    hipStream_t s = ctx->stream;
    {
        // fp4 grids of >= 256 big tiles: packed operands expanded in
        // registers (MELA_MFMA_EXPANDED=1 keeps the expanded-operand kernel)
        const size_t t256 = ((m + XROWS - 1) / XROWS) * ((n + XROWS - 1) / XROWS);
        static const bool expanded = [] {
            const char *e = std::getenv("MELA_MFMA_EXPANDED");
            return e && *e && *e != '0';
        }();
        // (k < 2^23: the direct kernel's accumulators start at 2^23)
        if (fp4 && !expanded && t256 >= (size_t)ctx->cu_count && k < ((size_t)1 << 23))
            return matmul_mfma_packed(ctx, A, lda, B, ldb, C, ldc, m, k, n, gate, count, sparse_path);
    }
    const size_t kw = words64(k);
    const size_t xbytes = kw * (fp4 ? 32 : 64);  // expanded bytes per row
    const size_t pitch = (xbytes + KPAD - 1) / KPAD * KPAD;
    const size_t mp = (m + XROWS - 1) / XROWS * XROWS, np = (n + XROWS - 1) / XROWS * XROWS;
    void *pa = nullptr, *pb = nullptr;
    mela_status st;
    if ((st = scratch_get(ctx, SLOT_XA, mp * pitch, &pa)) != MELA_OK) return st;
    if ((st = scratch_get(ctx, SLOT_XB, np * pitch, &pb)) != MELA_OK) return st;
    if (count) {
        // expanded-operand paths: a separate count, then the row-XOR product
        mela_status cs = launch_popcount(ctx, A, lda, m, kw, count);
        if (cs == MELA_OK && sparse_path) cs = sparse_path(gate);
        if (cs != MELA_OK) return cs;
    }
    const size_t ca = mp * (pitch / 32);
    const size_t kwp = pitch / (fp4 ? 32 : 64);  // 64-bit k words the padded pitch holds
    if (np / 64 > 65535 || (kwp + 3) / 4 > 0x7fffffffu)
        return set_error(MELA_ERR_INVALID, "too many columns for the matrix-core product");
    // 256 x 256 tiles (stage-tiled operands) once they fill every CU, else
    // 128 x 128 tiles, two per CU (row-major operands)
    const size_t tx2 = mp / 256, ty2 = np / 256;
    const bool ring = tx2 * ty2 >= (size_t)ctx->cu_count;
    // B^T: one 64 x 64 bit block per wave, two for stage-tiled fp4 (kwp is a
    // multiple of 8)
    const size_t kbw = ring && fp4 ? 2 : 1;
    const dim3 gt((unsigned)((kwp + 4 * kbw - 1) / (4 * kbw)), (unsigned)(np / 64));
    // stage-tiled A: 64 rows x 4 stage segments per workgroup (pitch is a
    // multiple of 4 segments)
    const dim3 gx((unsigned)(pitch / RING_STAGE / 4), (unsigned)(mp / 64));
    if (ring && mp / 64 > 65535)
        return set_error(MELA_ERR_INVALID, "too many rows for the matrix-core product");
    if (fp4) {
        if (ring)
            hipLaunchKernelGGL(k_expand_tiled<true>, gx, dim3(256), 0, s, A,
                               lda, m, kw, pitch, (uint8_t *)pa, gate);
        else
            hipLaunchKernelGGL(k_expand<true>, dim3((unsigned)((ca + 255) / 256)), dim3(256), 0, s,
                               A, lda, m, kw, mp, pitch, (uint8_t *)pa, gate);
        if (ring)
            hipLaunchKernelGGL((k_transpose_expand<true, 2>), gt, dim3(256), 0, s, B, ldb, k, n, kwp,
                               (uint8_t *)pb, pitch, true, gate);
        else
            hipLaunchKernelGGL((k_transpose_expand<true, 1>), gt, dim3(256), 0, s, B, ldb, k, n, kwp,
                               (uint8_t *)pb, pitch, false, gate);
    } else {
        if (ring)
            hipLaunchKernelGGL(k_expand_tiled<false>, gx, dim3(256), 0, s,
                               A, lda, m, kw, pitch, (uint8_t *)pa, gate);
        else
            hipLaunchKernelGGL(k_expand<false>, dim3((unsigned)((ca + 255) / 256)), dim3(256), 0, s,
                               A, lda, m, kw, mp, pitch, (uint8_t *)pa, gate);
        hipLaunchKernelGGL((k_transpose_expand<false, 1>), gt, dim3(256), 0, s, B, ldb, k, n, kwp,
                           (uint8_t *)pb, pitch, ring, gate);
    }
    MELA_HIP(hipGetLastError());
    const uint8_t *xa = (const uint8_t *)pa, *xb = (const uint8_t *)pb;
    prof_mark(ctx, true);
    if (ring) {
        if (tx2 * ty2 > (size_t)INT32_MAX) return set_error(MELA_ERR_INVALID, "too many tiles for one launch");
        const dim3 grid((unsigned)(tx2 * ty2));
        if (fp4)
            hipLaunchKernelGGL(k_mfma_ring<true>, grid, dim3(256), 0, s, xa, xb, pitch, C, ldc, (int)m, (int)n, (int)tx2, (int)ty2, gate);
        else
            hipLaunchKernelGGL(k_mfma_ring<false>, grid, dim3(256), 0, s, xa, xb, pitch, C, ldc, (int)m, (int)n, (int)tx2, (int)ty2, gate);
    } else {
        const size_t tx1 = (m + 127) / 128, ty1 = (n + 127) / 128;
        const dim3 grid((unsigned)(tx1 * ty1));
        if (fp4)
            hipLaunchKernelGGL(k_mfma_small<true>, grid, dim3(256), 0, s, xa, xb, pitch, C, ldc, (int)m, (int)n, (int)tx1, (int)ty1, gate);
        else
            hipLaunchKernelGGL(k_mfma_small<false>, grid, dim3(256), 0, s, xa, xb, pitch, C, ldc, (int)m, (int)n, (int)tx1, (int)ty1, gate);
    }
    prof_mark(ctx, false);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

}  // namespace mela
