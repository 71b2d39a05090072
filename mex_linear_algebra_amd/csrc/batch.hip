// Batched rank / RREF / canonical null basis for thousands of uneven matrices
// in one persistent launch (no reference counterpart: the reference handles
// one matrix per MEX call, e.g. null_gf2.c:142).
//
// Every workgroup loops: take the next matrix from a device queue (largest
// first), copy it into its working slot, eliminate it panel by panel entirely
// inside the workgroup (panel select -> pivot rows -> table or direct XOR
// sweep -> earlier pivot rows), then write its rank and its packed null
// basis.  Per-matrix mode (chosen by the autotuner): GE (XOR each selected
// pivot row) or M4RI with k-bit tables, k = 4 (16 table lookups per row
// word) or, in the kernel with 8-word tiles, k = 8 (8 lookups per row word,
// 8 x 256-entry tables per tile word); the update tile width (8, 16 or 32
// words) is chosen by the autotuner per batch.
//
// Per panel and tile of TW words: the selected rows' words are staged in LDS,
// the tile's pivot rows are formed from 4-bit tables of them (and stored to
// Q), the matrix's M4RI tables are built from those, and the sweep runs over
// the panel's list of rows with a nonzero multiplier: each lane takes two
// consecutive words of a row (one 16-byte table read serves both), several
// row passes per wave have their loads in flight together, and the lookups
// go table by table.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "elim_device.h"

namespace mela {

struct BatchMat {
    uint64_t a_off;  // words: input A (m x lda) in the input pointer table
    uint64_t w_off;  // words: working copy (m x wc) in the work arena
    uint64_t q_off;  // words: pivot rows (min(m,n) x wc) in the work arena
    uint64_t z_off;  // words: packed null basis (n x ldz) in the Z arena
    uint32_t m, n;
    uint32_t lda, ldz;
    uint32_t mode;    // 0 = GE, 1 = M4RI k = 4, 2 = M4RI k = 8 (8-word tiles; k = 4 otherwise)
    uint32_t groups;  // workgroups eliminating the matrix together (1..MAX_GROUP)
};

namespace {

constexpr int MAX_M = 4096;
constexpr int MAX_N = 8192;
constexpr int MAX_WC = MAX_N / 64;
constexpr int SW_PASSES = 16;    // row passes in flight per wave, before the per-kernel cut
constexpr int MAX_GROUP = 8;     // workgroups per matrix at most
// a group member that waits this long (wall clock ticks, 100 MHz: 0.5 s)
// for its partners gives up and flags the batch as failed
constexpr uint64_t GROUP_WAIT_TICKS = 50000000ull;

// The 8-bit tables of a tile: 8 tables x 256 entries x 8 words = 128 KiB,
// so one workgroup per CU; that kernel runs 512 threads per workgroup (two
// waves per SIMD, as the 256-thread kernels with 16- and 32-word tiles get
// from two resident workgroups).
constexpr int K8_TW = 8;
constexpr int T8_N = 8;
constexpr int batch_threads(int tw) { return tw == K8_TW ? 512 : 256; }

template <int TW>
struct BatchLDS {
    dev::SelectLDS<batch_threads(TW)> sel;
    uint64_t comb[64];
    int selrow[64];
    uint16_t act[MAX_M];
    uint16_t upd[MAX_M];     // sweep list of the panel (rows with a nonzero multiplier)
    uint64_t pmask[MAX_WC];  // pivot column bitmap
    int pbefore[MAX_WC];     // pivots in columns below 64*w
    uint64_t pw[64];         // word g of the panel's pivot rows, by pivot bit
    // One region, used in turn per tile: the selected rows' words, their
    // tables, the pivot rows' tables (M4RI) or the pivot rows (GE); the null
    // extraction's column word tables reuse it after the elimination.
    // 16-byte aligned: the sweep reads two words of a table entry at once
    union alignas(16) {
        uint64_t S[64][TW];      // selected rows' words of the tile
        uint64_t T[16][16][TW];  // 4-bit tables (of the selected rows, then of the pivot rows)
        uint64_t P[64][TW];      // pivot rows of the tile, by pivot bit (GE sweep, 8-bit tables)
        // 8-bit tables of the pivot rows (8-word tiles only): table t,
        // entry v = XOR of pivot rows 8t + b over the set bits b of v
        uint64_t T8[TW == K8_TW ? T8_N : 1][TW == K8_TW ? 256 : 1][TW];
        struct {
            uint64_t fm[MAX_WC];     // free-column mask of each column word
            uint64_t mv[MAX_WC][6];  // its compress masks
            int fb[MAX_WC];          // free columns before the word
            int ws[MAX_WC + 1];      // column word holding free column 64 j
        } nl;
    } u;
    int s_idx;
    int s_r, s_nact;
    int s_ok;
};

template <int TW>
__device__ __forceinline__ int pivots_before(const BatchLDS<TW> &L, int p) {
    const int w = p >> 6;
    return L.pbefore[w] + dev::popc64(L.pmask[w] & ((1ull << (p & 63)) - 1ull));
}

// Barrier of the G workgroups that eliminate one matrix together.  *bar
// counts arrivals (zeroed before the launch, never reset); the phase-th
// barrier waits for G * phase of them.  Every wave releases its stores to
// device scope before arriving (L2 write-back, the members may sit on
// different XCDs) and acquires after the wait (L1/L2 invalidate).  A member
// that waits longer than GROUP_WAIT_TICKS flags *err and returns false, and
// once *err is set every group barrier returns false at once, so the launch
// drains quickly and every wave ends; the host then reports the batch as
// failed.  With G == 1 it is a workgroup barrier.
__device__ __forceinline__ bool group_sync(unsigned *bar, int G, unsigned phase, int *err, int &s_ok) {
    if (G == 1) {
        __syncthreads();
        return true;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned target = (unsigned)G * phase;
        const uint64_t t0 = wall_clock64();
        int ok = 1;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > GROUP_WAIT_TICKS) {
                atomicOr(err, 1);
                ok = 0;
                break;
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return s_ok != 0;
}

// A value every lane of the wave holds alike, moved into scalar registers
// (each 32-bit word read from the first active lane).  A work source reads
// its descriptors through struct members, which carry no no-alias promise,
// so the compiler loads them into vector registers; without this every
// matrix parameter and everything derived from it would occupy VGPRs.
template <class T>
__device__ __forceinline__ T wave_uniform(const T &v) {
    static_assert(sizeof(T) % 4 == 0, "32-bit words");
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)w[i]);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

// The matrix one workgroup works on next, as a work source hands it out:
// member mk of the matrix's d.groups workgroups, its group barrier, input,
// working memory (rows at work + d.w_off, pivot rows at work + d.q_off),
// null-basis arena, and the source's index of the matrix (its rank slot).
struct BatchItem {
    BatchMat d;
    const uint64_t *A;
    uint64_t *work;
    uint64_t *zarena;
    unsigned *bar;
    long long seq;  // the serving ring's sequence number (ServeSource)
    int idx;
    int mk;
    int want_null;
};

// The elimination loop, one kernel for every work source: each workgroup
// takes the next matrix from src.next() (called by every thread; false ends
// the loop), copies it in, eliminates it panel by panel, writes its rank and
// (want_null) its packed canonical null basis into Z, and hands it back to
// src.finish(); src.rank() stores the rank.  A group barrier that gave up skips the rest of the matrix
// (the batch then fails as a whole; only a batch source has groups).
template <int TW, class Src>
__global__ void __launch_bounds__(batch_threads(TW))
k_rref_loop(const Src src_arg) {
    constexpr int BT = batch_threads(TW);  // threads per workgroup
    constexpr int BT_WAVES = BT / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    BatchLDS<TW> &L = *reinterpret_cast<BatchLDS<TW> *>(smem);
    Src src = src_arg;
    int *err = src.err();
    for (;;) {
    BatchItem it;
    if (!src.next(L, it)) break;
    const BatchMat d = it.d;
    const int mk = it.mk;
    unsigned *bar = it.bar;
    const uint64_t *A = it.A;
    uint64_t *work = it.work;
    uint64_t *zarena = it.zarena;
    const int want_null = it.want_null;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int G = (int)d.groups;
    unsigned phase = 0;
    const int m = (int)d.m, n = (int)d.n;
    const int wc = (n + 63) >> 6;
    const int rmax = m < n ? m : n;
    // table bits of the sweep: 0 (GE), 4, or 8 (8-word tiles only)
    // (the host passes mode 2 to the 8-word tile kernel only)
    const int kb = d.mode == 1 ? 4 : (TW == K8_TW && d.mode == 2 ? 8 : 0);
    uint64_t *M = work + d.w_off;
    uint64_t *Q = work + d.q_off;

    {
        // member mk copies its share of the words
        constexpr int CU = 8;  // words per thread with loads in flight together
        const int share = (m * wc + G - 1) / G;
        const int mw = min(m * wc, (mk + 1) * share);
        for (int i0 = mk * share + tid; i0 < mw; i0 += BT * CU) {
            uint64_t v[CU];
#pragma unroll
            for (int k = 0; k < CU; ++k) {
                const int i = i0 + k * BT < mw ? i0 + k * BT : i0;
                const int r = i / wc, w = i - r * wc;
                v[k] = A[(size_t)r * d.lda + w];
            }
#pragma unroll
            for (int k = 0; k < CU; ++k)
                if (i0 + k * BT < mw) M[i0 + k * BT] = v[k];
        }
    }
    for (int i = tid; i < m; i += BT) L.act[i] = (uint16_t)i;
    for (int i = tid; i < wc; i += BT) L.pmask[i] = 0;
    if (tid == 0) {
        L.s_r = 0;
        L.s_nact = m;
    }
    if (!group_sync(bar, G, ++phase, err, L.s_ok)) continue;

    // Members run the same panel selection on the same words (identical
    // results); the tiles of each panel are dealt out round robin; one
    // group barrier per panel.  Member 0 owns tile 0 (word g), so it
    // updates word g of the earlier pivot rows, after the barrier (when
    // every member has read it).
    bool dropped = false;
    for (int g = 0; g < wc; ++g) {
        const int n_act = L.s_nact, r0 = L.s_r;
        if (n_act == 0 || r0 >= rmax) break;
        auto rowfn = [&](int q) { return M + (size_t)q * wc; };
        const dev::PanelOut po =
            dev::wg_panel_select<BT>(L.sel, L.act, n_act, g, rowfn, L.comb, L.selrow);
        const uint64_t pm = po.pm;
        if (pm == 0) continue;  // no pivot in this panel (uniform)
        const int rho = po.rho;
        if (tid == 0) {
            L.pmask[g] = pm;
            L.s_r = r0 + rho;
            L.s_nact = po.n_act_new;
        }
        // the new pivot rows are zero left of their panel
        for (int e = mk == 0 ? tid : rho * g; e < rho * g; e += BT) {
            const int j = e / g, w = e - j * g;
            Q[(size_t)(r0 + j) * wc + w] = 0ull;
        }
        // active rows and (for the null basis) earlier pivot rows
        const int nact = po.n_act_new;
        const int total = nact + (want_null ? r0 : 0);
        // rows of sweep positions: active rows by row id, earlier pivot
        // rows (Q) by 0x8000 | index
        auto sweep_row = [&](uint32_t code) -> uint64_t * {
            return code & 0x8000u ? Q + (size_t)(code & 0x7fffu) * wc : M + (size_t)code * wc;
        };
        // list of the positions whose multiplier (word g & pm) is nonzero
        int nupd = 0;
        {
            constexpr int CK = 4;  // positions per thread per round, loads together
            for (int b0 = 0; b0 < total; b0 += BT * CK) {
                uint32_t code[CK];
                bool f[CK];
#pragma unroll
                for (int k = 0; k < CK; ++k) {
                    const int q = b0 + k * BT + tid;
                    const int qa = q < nact ? q : 0;
                    code[k] = q < nact ? (uint32_t)L.act[qa] : 0x8000u | (uint32_t)(q - nact);
                    if (q >= total) code[k] = 0;
                }
                uint64_t xv[CK];
#pragma unroll
                for (int k = 0; k < CK; ++k) xv[k] = sweep_row(code[k])[g];
                // wave-local slots, k-major: entry k of lane l goes to
                // (entries of rounds < k) + (lanes < l with entry k)
                int wsum = 0, off[CK];
                const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
                for (int k = 0; k < CK; ++k) {
                    f[k] = b0 + k * BT + tid < total && (xv[k] & pm) != 0ull;
                    const uint64_t bal = __ballot(f[k]);
                    off[k] = wsum + dev::popc64(bal & below);
                    wsum += dev::popc64(bal);
                }
                if (lane == 0) L.sel.wcnt[wave] = wsum;
                __syncthreads();
                int base = nupd, all = 0;
#pragma unroll
                for (int v = 0; v < BT_WAVES; ++v) {
                    const int c = L.sel.wcnt[v];
                    if (v < wave) base += c;
                    all += c;
                }
#pragma unroll
                for (int k = 0; k < CK; ++k)
                    if (f[k]) L.upd[base + off[k]] = (uint16_t)code[k];
                nupd += all;
                __syncthreads();
            }
        }
        // the sweep's lanes: TW / 2 per row, two consecutive words each
        // (one 16-byte table read serves both); RPP rows per wave pass
        constexpr int LPR = TW / 2;
        constexpr int RPP = 64 / LPR;
        // row passes per wave with loads in flight together (the 512-thread
        // kernel has twice the waves per SIMD and half the registers)
        constexpr int SWP = TW == K8_TW ? SW_PASSES / 4 : SW_PASSES / 2;
        const int sub = lane / LPR, wp = lane % LPR;
        for (int w0 = g + mk * TW; w0 < wc; w0 += G * TW) {
            // selected rows' words of the tile (loads issued together;
            // slots past rho and words past wc read row 0 and store 0)
            {
                constexpr int SE = 64 * TW / BT;
                uint64_t sv[SE];
#pragma unroll
                for (int k = 0; k < SE; ++k) {
                    const int e = tid + k * BT, slot = e / TW, ww = e - slot * TW;
                    const bool ok = slot < rho && w0 + ww < wc;
                    sv[k] = M[ok ? (size_t)L.selrow[slot] * wc + w0 + ww : 0];
                }
#pragma unroll
                for (int k = 0; k < SE; ++k) {
                    const int e = tid + k * BT, slot = e / TW, ww = e - slot * TW;
                    L.u.S[slot][ww] = (slot < rho && w0 + ww < wc) ? sv[k] : 0ull;
                }
            }
            __syncthreads();
            // Thread (t, ww) owns table t of tile word ww: first the 4-bit
            // tables of the selected rows 4t..4t+3, from which it forms
            // the pivot rows 4t..4t+3 (pivot row c = XOR of the selected
            // rows in comb[c], 16 lookups), then the tables of those.
            constexpr int TPT = (16 * TW + BT - 1) / BT;  // tables per thread
            uint64_t q[TPT][4];
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                const int e = tid + k * BT, t = e / TW, ww = e - t * TW;
#pragma unroll
                for (int b = 0; b < 4; ++b) q[k][b] = e < 16 * TW ? L.u.S[4 * t + b][ww] : 0ull;
            }
            __syncthreads();
            auto build_tables = [&]() {
#pragma unroll
                for (int k = 0; k < TPT; ++k) {
                    const int e = tid + k * BT, t = e / TW, ww = e - t * TW;
                    if (e >= 16 * TW) continue;
                    uint64_t cur = 0;
                    L.u.T[t][0][ww] = 0;
#pragma unroll
                    for (int v = 1; v < 16; ++v) {
                        cur ^= q[k][__builtin_ctz(v)];
                        L.u.T[t][v ^ (v >> 1)][ww] = cur;  // Gray-code order
                    }
                }
            };
            build_tables();
            __syncthreads();
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                const int e = tid + k * BT, t = e / TW, ww = e - t * TW;
                const int w = w0 + ww;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int c = 4 * t + b;
                    const bool pc = e < 16 * TW && ((pm >> c) & 1ull);
                    const uint64_t cb = pc ? L.comb[c] : 0ull;
                    uint64_t v = 0;
#pragma unroll
                    for (int u = 0; u < 16; ++u) v ^= L.u.T[u][(cb >> (4 * u)) & 15u][e < 16 * TW ? ww : 0];
                    q[k][b] = v;
                    if (pc && w < wc) Q[(size_t)(r0 + dev::popc64(pm & ((1ull << c) - 1ull))) * wc + w] = v;
                    if (pc && w == g) L.pw[c] = v;
                }
            }
            __syncthreads();
            if (kb == 4) {
                build_tables();
            } else {
#pragma unroll
                for (int k = 0; k < TPT; ++k) {
                    const int e = tid + k * BT, t = e / TW, ww = e - t * TW;
                    if (e >= 16 * TW) continue;
#pragma unroll
                    for (int b = 0; b < 4; ++b) L.u.P[4 * t + b][ww] = q[k][b];
                }
                if constexpr (TW == K8_TW) {
                    if (kb == 8) {
                        // thread (t8, ch, ww) builds entries v = 8 j + ch of
                        // table t8 for tile word ww, j in Gray-code order (a
                        // wave's stores cover 512 consecutive bytes)
                        static_assert(BT == T8_N * 8 * K8_TW, "one thread per eighth of a table");
                        __syncthreads();
                        const int ww = tid % K8_TW, ch = (tid / K8_TW) & 7, t8 = tid / (8 * K8_TW);
                        uint64_t pr[8];
#pragma unroll
                        for (int b = 0; b < 8; ++b) pr[b] = L.u.P[8 * t8 + b][ww];
                        __syncthreads();
                        uint64_t cur = ((ch & 1) ? pr[0] : 0ull) ^ ((ch & 2) ? pr[1] : 0ull) ^
                                       ((ch & 4) ? pr[2] : 0ull);
                        L.u.T8[t8][ch][ww] = cur;
#pragma unroll
                        for (int j = 1; j < 32; ++j) {
                            cur ^= pr[3 + __builtin_ctz(j)];
                            L.u.T8[t8][8 * (j ^ (j >> 1)) + ch][ww] = cur;
                        }
                    }
                }
            }
            __syncthreads();
            // sweep words > g of the tile, words w and w + 1 per lane; SWP
            // passes' loads together
            const int w = w0 + 2 * wp;
            const bool ok0 = w < wc && w > g, ok1 = w + 1 < wc && w + 1 > g;
            const int ws0 = ok0 ? w : g, ws1 = ok1 ? w + 1 : g;
#pragma unroll 1
            for (int i0 = wave * RPP + sub; i0 < nupd; i0 += BT_WAVES * RPP * SWP) {
                // every pass loads unconditionally (entries past the end
                // re-read row 0 of M) so all loads are in flight together
                uint64_t *rp[SWP];
                uint64_t x[SWP], a0[SWP], a1[SWP];
#pragma unroll
                for (int j = 0; j < SWP; ++j) {
                    const int i = i0 + j * BT_WAVES * RPP;
                    rp[j] = i < nupd ? sweep_row(L.upd[i]) : M;
                }
#pragma unroll
                for (int j = 0; j < SWP; ++j) x[j] = rp[j][g];
#pragma unroll
                for (int j = 0; j < SWP; ++j) {
                    a0[j] = rp[j][ws0];
                    a1[j] = rp[j][ws1];
                }
#pragma unroll
                for (int j = 0; j < SWP; ++j) {
                    const int i = i0 + j * BT_WAVES * RPP;
                    x[j] = (i < nupd && (ok0 || ok1)) ? x[j] & pm : 0ull;
                }
                auto acc2 = [&](int j, const uint64_t *e) {
                    const uint4 q = *reinterpret_cast<const uint4 *>(e);
                    a0[j] ^= ((uint64_t)q.y << 32) | q.x;
                    a1[j] ^= ((uint64_t)q.w << 32) | q.z;
                };
                if (kb == 4) {
                    // table-major: one lookup of every pass per table, so
                    // the passes' LDS reads overlap (a zero multiplier
                    // reads the zero entry)
#pragma unroll
                    for (int t = 0; t < 16; ++t)
#pragma unroll
                        for (int j = 0; j < SWP; ++j) acc2(j, &L.u.T[t][(x[j] >> (4 * t)) & 15u][2 * wp]);
                } else if (kb == 8) {
                    if constexpr (TW == K8_TW) {
#pragma unroll
                        for (int t = 0; t < T8_N; ++t)
#pragma unroll
                            for (int j = 0; j < SWP; ++j) acc2(j, &L.u.T8[t][(x[j] >> (8 * t)) & 255u][2 * wp]);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < SWP; ++j)
                        for (uint64_t xx = x[j]; xx; xx &= xx - 1) acc2(j, &L.u.P[dev::ffs64(xx)][2 * wp]);
                }
#pragma unroll
                for (int j = 0; j < SWP; ++j)
                    if (x[j]) {
                        if (ok0) rp[j][w] = a0[j];
                        if (ok1) rp[j][w + 1] = a1[j];
                    }
            }
            __syncthreads();
        }
        if (!group_sync(bar, G, ++phase, err, L.s_ok)) {
            dropped = true;
            break;
        }
        // word g of the earlier pivot rows (every member has read it)
        if (want_null && mk == 0) {
            constexpr int PU = 4;  // rows per thread with loads in flight together
            for (int q0 = tid; q0 < r0; q0 += BT * PU) {
                uint64_t xv[PU];
#pragma unroll
                for (int k = 0; k < PU; ++k) {
                    const int q = q0 + k * BT < r0 ? q0 + k * BT : q0;
                    xv[k] = Q[(size_t)q * wc + g];
                }
#pragma unroll
                for (int k = 0; k < PU; ++k) {
                    const int q = q0 + k * BT;
                    if (q >= r0) continue;
                    uint64_t a = xv[k];
                    for (uint64_t h = xv[k] & pm; h; h &= h - 1) a ^= L.pw[dev::ffs64(h)];
                    Q[(size_t)q * wc + g] = a;
                }
            }
        }
        __syncthreads();
    }

    if (dropped) continue;
    // the last panel's pivot row words before any member extracts
    if (!group_sync(bar, G, ++phase, err, L.s_ok)) continue;
    const int r = L.s_r;
    if (tid == 0 && mk == 0) src.rank(it, (uint32_t)r);
    if (want_null) {
        const int f = n - r;
        // prefix pivot counts per column word
        if (tid == 0) {
            int acc = 0;
            for (int w = 0; w < wc; ++w) {
                L.pbefore[w] = acc;
                acc += dev::popc64(L.pmask[w]);
            }
        }
        __syncthreads();
        // per column word: free mask, compress masks, free columns before
        // it; ws[j] = the column word that holds free column 64 j
        for (int w = tid; w < wc; w += BT) {
            const int valid = n - 64 * w < 64 ? n - 64 * w : 64;
            const uint64_t fm = ~L.pmask[w] & (valid == 64 ? ~0ull : ((1ull << valid) - 1ull));
            dev::compress_masks(fm, L.u.nl.mv[w]);
            L.u.nl.fm[w] = fm;
            const int fb = 64 * w - L.pbefore[w];
            L.u.nl.fb[w] = fb;
            const int j = (fb + 63) >> 6;
            if (64 * j < fb + dev::popc64(fm)) L.u.nl.ws[j] = w;
        }
        __syncthreads();
        uint64_t *Z = zarena + d.z_off;
        const int ldz = (int)d.ldz;
        const int fw = (f + 63) >> 6;
        // Z row p: a pivot column's row is its pivot row's free columns
        // (word jw gathers the compressed column words ws[jw]..ws[jw+1]),
        // a free column's row is its unit vector.  Four words per thread
        // per step, their row-word loads issued together.
        constexpr int NZ = 4;
        const int zshare = (n * ldz + G - 1) / G;
        const int zend = min(n * ldz, (mk + 1) * zshare);
        for (int i0 = mk * zshare + tid; i0 < zend; i0 += BT * NZ) {
            const uint64_t *qrow[NZ];
            int lo[NZ], wa[NZ], wb[NZ];
            uint64_t word[NZ];
#pragma unroll
            for (int k = 0; k < NZ; ++k) {
                const int i = i0 + k * BT;
                const int p = i / ldz, jw = i - p * ldz;
                qrow[k] = nullptr;
                word[k] = 0;
                wa[k] = 0;
                wb[k] = -1;
                lo[k] = 64 * jw;
                if (i < zend && jw < fw) {
                    const int pb = pivots_before(L, p);
                    if ((L.pmask[p >> 6] >> (p & 63)) & 1ull) {
                        qrow[k] = Q + (size_t)pb * wc;
                        wa[k] = L.u.nl.ws[jw];
                        wb[k] = jw + 1 < fw ? L.u.nl.ws[jw + 1] : wc - 1;
                    } else {
                        const int fi = p - pb;
                        if ((fi >> 6) == jw) word[k] = 1ull << (fi & 63);
                    }
                }
            }
            // the first three column words of each output word, loaded
            // together (unit rows and words past the end re-read Q)
            uint64_t x0[NZ], x1[NZ], x2[NZ];
#pragma unroll
            for (int k = 0; k < NZ; ++k) {
                const uint64_t *src = qrow[k] ? qrow[k] : Q;
                x0[k] = src[wa[k] < wc ? wa[k] : wc - 1];
                x1[k] = src[wa[k] + 1 < wc ? wa[k] + 1 : wc - 1];
                x2[k] = src[wa[k] + 2 < wc ? wa[k] + 2 : wc - 1];
            }
#pragma unroll
            for (int k = 0; k < NZ; ++k) {
                // column word w holds free columns fb[w] .. fb[w] + popc - 1
                auto put = [&](int w, uint64_t xv) {
                    const int fb = L.u.nl.fb[w];
                    if (fb >= lo[k] + 64) return;
                    const uint64_t c = dev::compress_apply(xv, L.u.nl.fm[w], L.u.nl.mv[w]);
                    const int sh = fb - lo[k];
                    word[k] |= sh >= 0 ? c << sh : c >> (-sh);
                };
                const int nw = wb[k] - wa[k] + 1;
                if (nw > 0) put(wa[k], x0[k]);
                if (nw > 1) put(wa[k] + 1, x1[k]);
                if (nw > 2) put(wa[k] + 2, x2[k]);
                for (int w = wa[k] + 3; w <= wb[k]; ++w) put(w, qrow[k][w]);
                const int i = i0 + k * BT;
                if (i < zend) Z[i] = word[k];
            }
        }
    }
    src.finish(it);
    __syncthreads();
    }
}

// Work source of a batch launch.  Slot s of the schedule is member
// slots[s].y of queue entry slots[s].x (matrix order[x]); a matrix's G slots
// are consecutive, so at most one matrix at a time is partly staffed and
// every member is taken by a running workgroup.
struct BatchSource {
    const BatchMat *__restrict__ mats;
    const int *__restrict__ order;
    const int2 *__restrict__ slots;
    int *__restrict__ counter;  // [0] next slot, [1] error flag
    unsigned *__restrict__ gbar;
    const uint64_t *const *__restrict__ inputs;
    uint64_t *__restrict__ work;
    uint64_t *__restrict__ zarena;
    uint32_t *__restrict__ ranks;
    int nslots;
    int want_null;

    __device__ int *err() const { return counter + 1; }
    template <class LDS>
    __device__ bool next(LDS &L, BatchItem &it) {
        if (threadIdx.x == 0) L.s_idx = atomicAdd(counter, 1);
        __syncthreads();
        const int si = __builtin_amdgcn_readfirstlane(L.s_idx);
        if (si >= nslots) return false;
        const int2 sl = wave_uniform(slots[si]);
        const int mi = __builtin_amdgcn_readfirstlane(order[sl.x]);
        it.d = wave_uniform(mats[mi]);
        it.A = wave_uniform(inputs[mi]);
        it.work = work;
        it.zarena = zarena;
        it.bar = gbar + sl.x;
        it.seq = 0;
        it.idx = mi;
        it.mk = sl.y;
        it.want_null = want_null;
        return true;
    }
    __device__ void rank(const BatchItem &it, uint32_t r) { ranks[it.idx] = r; }
    __device__ void finish(const BatchItem &) {}
};

// Work source of the device-resident serving loop: each workgroup takes
// published request slots (compare-and-swap on head, in device memory, while
// head < tail), one workgroup per matrix, and reports each one; it ends when
// the host sets the stop bit of tail, when no slot was published for
// idle_ticks, or, between requests, once the launch is life_ticks old (the
// host starts the next instance at once when work remains: a kernel that
// never ended would hold up every stream that shares its hardware queue).
// An idle workgroup polls the host's tail every ~0.3 us at first and every
// ~4 us after 50 us without work.
struct ServeSource {
    const ServeSlot *__restrict__ slots;
    ServeDone *__restrict__ done;
    const uint64_t *__restrict__ tail;
    uint64_t *__restrict__ head;
    uint64_t idle_ticks;
    uint64_t life_ticks;
    uint64_t t_start;  // set by the kernel's first next()
    uint64_t t_idle;
    uint32_t cap;

    __device__ int *err() const { return nullptr; }  // one workgroup per matrix
    template <class LDS>
    __device__ bool next(LDS &, BatchItem &it) {
        __shared__ long long s_seq;
        __shared__ ServeSlot s_slot;
        if (t_start == 0) t_start = t_idle = wall_clock64();
        if (threadIdx.x == 0) {
            long long got = -1;
            for (;;) {
                if (wall_clock64() - t_start > life_ticks) break;
                const uint64_t t = __hip_atomic_load(tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
                if (t & SERVE_STOP) break;
                const uint64_t h = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (h < t) {
                    uint64_t expect = h;
                    if (__hip_atomic_compare_exchange_strong(head, &expect, h + 1, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        got = (long long)h;
                        break;
                    }
                    continue;
                }
                const uint64_t idle = wall_clock64() - t_idle;
                if (idle > idle_ticks) break;
                if (idle > 5000) {  // 50 us
                    __builtin_amdgcn_s_sleep(127);
                    __builtin_amdgcn_s_sleep(127);
                } else {
                    __builtin_amdgcn_s_sleep(8);
                }
            }
            if (got >= 0) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // the slot was written before tail
                s_slot = slots[(uint64_t)got % cap];  // one read of host memory per request
            }
            s_seq = got;
        }
        __syncthreads();
        const long long seq = wave_uniform(s_seq);
        if (seq < 0) return false;
        const ServeSlot sl = wave_uniform(s_slot);
        const uint32_t wc = (sl.n + 63) / 64;
        it.d = BatchMat{};
        it.d.m = sl.m;
        it.d.n = sl.n;
        it.d.lda = sl.lda;
        it.d.ldz = wc;
        it.d.w_off = 0;
        it.d.q_off = (uint64_t)sl.m * wc;
        it.d.z_off = 0;
        it.d.mode = sl.mode;
        it.d.groups = 1;
        it.A = sl.A;
        it.work = sl.work;
        it.zarena = sl.Z;
        it.bar = nullptr;
        it.seq = seq;
        it.idx = (int)((uint64_t)seq % cap);
        it.mk = 0;
        it.want_null = sl.op == MELA_OP_NULL ? 1 : 0;
        return true;
    }
    __device__ void rank(const BatchItem &it, uint32_t r) { done[it.idx].rank = r; }
    // the rank, the null basis and the working rows are written: make them
    // visible to the host and its copy engines, then report (the loop's
    // closing barrier keeps s_seq until every thread has read it)
    __device__ void finish(const BatchItem &it) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(&done[it.idx].seq1, (uint64_t)it.seq + 1, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        t_idle = wall_clock64();
    }
};

// queue counter, error flag and the group barrier counters start at zero
__global__ void k_reset_batch(int *counter, unsigned *gbar, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2) counter[i] = 0;
    if (i < count) gbar[i] = 0u;
}

// ones[i] += popcount of matrix i (one block per (matrix, row chunk))
__global__ void __launch_bounds__(256)
k_batch_popcount(const BatchMat *__restrict__ mats, const uint64_t *const *__restrict__ inputs,
                 unsigned long long *__restrict__ ones) {
    const BatchMat d = mats[blockIdx.x];
    const int wc = (int)((d.n + 63) >> 6);
    const uint64_t *A = inputs[blockIdx.x];
    unsigned long long acc = 0;
    for (int i = threadIdx.x; i < (int)d.m * wc; i += 256) {
        const int r = i / wc, w = i - r * wc;
        acc += (unsigned long long)__popcll(A[(size_t)r * d.lda + w]);
    }
    atomicAdd(&ones[blockIdx.x], acc);
}

}  // namespace

static_assert(sizeof(BatchLDS<K8_TW>) <= 160 * 1024, "8-word tile LDS exceeds a CU's 160 KiB");

size_t batch_lds_bytes(int tw) {
    return tw == 32 ? sizeof(BatchLDS<32>) : tw == K8_TW ? sizeof(BatchLDS<K8_TW>) : sizeof(BatchLDS<16>);
}

// index of a tile width (8, 16, 32 words) in per-width tables
int tw_slot(int tw) { return tw == K8_TW ? 0 : tw == 16 ? 1 : 2; }

// whether the device granted the 8-word tile kernel its LDS (set once)
bool g_k8_ok = true;

// Serving-loop workgroups: a quarter of the CUs.  The loop stays resident
// between requests, so the CUs it does not occupy stay free for the kernels
// of large requests on the worker streams (an MFMA product workgroup needs a
// whole CU's registers); 64 workgroups of <= 0.35 ms requests still take
// well over 10^5 requests/s.
int serve_loop_grid(mela_ctx *ctx) {
    const int cu = ctx->cu_count > 0 ? ctx->cu_count : 256;
    return cu / 4 > 8 ? cu / 4 : 8;
}

mela_status launch_serve_loop(hipStream_t s, int grid, ServeSlot *slots, ServeDone *done, const uint64_t *tail,
                              uint64_t *head, uint32_t cap, uint64_t idle_ticks, uint64_t life_ticks) {
    static std::mutex mu;
    static bool attr = false;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!attr) {
            MELA_HIP(hipFuncSetAttribute((const void *)k_rref_loop<16, ServeSource>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)batch_lds_bytes(16)));
            attr = true;
        }
    }
    ServeSource src{};
    src.slots = slots;
    src.done = done;
    src.tail = tail;
    src.head = head;
    src.idle_ticks = idle_ticks;
    src.life_ticks = life_ticks;
    src.cap = cap;
    hipLaunchKernelGGL((k_rref_loop<16, ServeSource>), dim3((unsigned)grid), dim3(batch_threads(16)),
                       batch_lds_bytes(16), s, src);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

}  // namespace mela

// ----------------------------------------------------------------- host side
struct mela_batch {
    mela_ctx *ctx = nullptr;
    size_t count = 0;
    int flags = 0;
    std::vector<mela::BatchMat> mats;
    std::vector<int> order;
    std::vector<const uint64_t *> inputs;
    mela::BatchMat *d_mats = nullptr;
    int *d_order = nullptr;
    int *d_counter = nullptr;  // [0] queue slot counter, [1] group wait error flag
    int2 *d_slots = nullptr;   // queue slots (entry, member), up to MAX_GROUP per matrix
    unsigned *d_gbar = nullptr;  // group barrier counter per queue entry
    int nslots = 0;
    const uint64_t **d_inputs = nullptr;
    uint64_t *d_work = nullptr;
    uint64_t *d_z = nullptr;
    uint32_t *d_ranks = nullptr;
    size_t work_words = 0, z_words = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int tuned = -1;  // index of the autotuned mode assignment
    int tw = 16;     // update tile width in words (8, 16 or 32, autotuned)
    int occ_tw[3] = {1, 1, 1};  // resident workgroups per CU for tile widths 8, 16 and 32
    bool ctx_scratch = false;  // arenas borrowed from the context (not freed)
};

namespace mela {
mela_status batch_launch_raw(mela_batch *b, hipStream_t s);
}

using namespace mela;

namespace mela {

const void *batch_kernel(int tw) {
    return tw == 32 ? (const void *)k_rref_loop<32, BatchSource>
                    : tw == K8_TW ? (const void *)k_rref_loop<K8_TW, BatchSource>
                                  : (const void *)k_rref_loop<16, BatchSource>;
}

// Persistent grid: every workgroup resident at once (the group barriers wait
// for partners that must be running), never more than the queue's slots.
int batch_grid(const mela_batch *b) {
    const int occ = b->occ_tw[tw_slot(b->tw)];
    return std::max(1, std::min(b->nslots, occ * b->ctx->cu_count));
}

mela_status batch_launch_raw(mela_batch *b, hipStream_t s) {
    const int nb = (int)b->mats.size();
    hipLaunchKernelGGL(k_reset_batch, dim3((unsigned)((std::max(nb, 2) + 255) / 256)), dim3(256), 0, s,
                       b->d_counter, b->d_gbar, nb);
    const int want_null = (b->flags & MELA_BATCH_NULL) ? 1 : 0;
    const int grid = batch_grid(b);
    BatchSource src{};
    src.mats = b->d_mats;
    src.order = b->d_order;
    src.slots = b->d_slots;
    src.counter = b->d_counter;
    src.gbar = b->d_gbar;
    src.inputs = b->d_inputs;
    src.work = b->d_work;
    src.zarena = b->d_z;
    src.ranks = b->d_ranks;
    src.nslots = b->nslots;
    src.want_null = want_null;
    const dim3 blk((unsigned)batch_threads(b->tw));
    if (b->tw == 32)
        hipLaunchKernelGGL((k_rref_loop<32, BatchSource>), dim3(grid), blk, batch_lds_bytes(32), s, src);
    else if (b->tw == K8_TW)
        hipLaunchKernelGGL((k_rref_loop<K8_TW, BatchSource>), dim3(grid), blk, batch_lds_bytes(K8_TW), s, src);
    else
        hipLaunchKernelGGL((k_rref_loop<16, BatchSource>), dim3(grid), blk, batch_lds_bytes(16), s, src);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

// Descriptors as the kernel of the batch's tile width reads them: mode 2
// (k = 8) runs as mode 1 (k = 4) unless the tiles are 8 words wide.
mela_status batch_upload_mats(mela_batch *b, hipStream_t s) {
    if (b->mats.empty()) return MELA_OK;
    std::vector<BatchMat> dm = b->mats;
    if (b->tw != K8_TW)
        for (BatchMat &d : dm)
            if (d.mode == 2) d.mode = 1;
    MELA_HIP(hipMemcpyAsync(b->d_mats, dm.data(), dm.size() * sizeof(BatchMat), hipMemcpyHostToDevice, s));
    MELA_HIP(hipStreamSynchronize(s));  // dm is freed on return
    return MELA_OK;
}

// Queue slots from the largest-first order and the per-matrix group sizes
// (a matrix's slots are consecutive), uploaded with the descriptors.
mela_status batch_upload_plan(mela_batch *b) {
    std::vector<int2> sl;
    sl.reserve(b->mats.size() * 2);
    for (size_t qi = 0; qi < b->order.size(); ++qi) {
        const int G = (int)b->mats[b->order[qi]].groups;
        for (int k = 0; k < G; ++k) sl.push_back(make_int2((int)qi, k));
    }
    b->nslots = (int)sl.size();
    hipStream_t s = b->ctx->stream;
    if (!b->mats.empty())
        MELA_HIP(hipMemcpyAsync(b->d_slots, sl.data(), sl.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    mela_status st = batch_upload_mats(b, s);
    if (st != MELA_OK) return st;
    // a captured graph holds the old grid size
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    return MELA_OK;
}

}  // namespace mela

namespace {

void free_batch_device(mela_batch *b) {
    if (!b->ctx_scratch) {
        void *ptrs[] = {b->d_mats,   b->d_order, b->d_counter, b->d_slots, b->d_gbar,
                        b->d_inputs, b->d_work,  b->d_z,       b->d_ranks};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
}

template <typename T>
mela_status dalloc(T **p, size_t count) {
    hipError_t e = hipMalloc((void **)p, (count ? count : 1) * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return set_error(MELA_ERR_NOMEM, "Out of memory.", "mela:memory");
    }
    return MELA_OK;
}

}  // namespace

extern "C" {

mela_status mela_batch_create(mela_ctx *ctx, const mela_bitmat *mats, size_t count, int flags,
                              mela_batch **out) {
    if (!ctx || !out || (count && !mats)) return set_error(MELA_ERR_INVALID, "null argument");
    *out = nullptr;
    // the context's stream and cached arenas are used below
    CtxLock lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    static std::mutex attr_mu;
    static bool attr_set = false;
    std::lock_guard<std::mutex> alk(attr_mu);
    if (!attr_set) {
        for (int tw : {16, 32})
            MELA_HIP(hipFuncSetAttribute(batch_kernel(tw), hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)batch_lds_bytes(tw)));
        // the 8-word tiles' LDS (about 157 KiB) is optional: a device that
        // does not grant it keeps the other tile widths
        if (hipFuncSetAttribute(batch_kernel(K8_TW), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)batch_lds_bytes(K8_TW)) != hipSuccess) {
            (void)hipGetLastError();
            g_k8_ok = false;
        }
        attr_set = true;
    }
    mela_batch *b = new mela_batch();
    b->ctx = ctx;
    b->count = count;
    b->flags = flags;
    size_t work = 0, z = 0;
    for (size_t i = 0; i < count; ++i) {
        const mela_bitmat &M = mats[i];
        if ((M.cols && M.ld < words64(M.cols)) || (M.rows && M.cols && !M.data)) {
            delete b;
            return set_error(MELA_ERR_INVALID, "bad matrix descriptor in batch");
        }
        if (M.rows > (size_t)MAX_M || M.cols > (size_t)MAX_N) {
            delete b;
            return set_error(MELA_ERR_INVALID,
                             "batch matrices are limited to 4096 x 8192; use the single-matrix "
                             "entry points for larger ones");
        }
        BatchMat d{};
        d.m = (uint32_t)M.rows;
        d.n = (uint32_t)M.cols;
        d.lda = (uint32_t)M.ld;
        const size_t wc = words64(M.cols);
        const size_t rmax = std::min(M.rows, M.cols);
        d.w_off = work;
        work += M.rows * wc;
        d.q_off = work;
        work += rmax * wc;
        d.ldz = (uint32_t)wc;
        d.z_off = z;
        z += M.cols * wc;
        d.mode = (flags & MELA_BATCH_GE) ? 0 : 1;
        d.groups = 1;
        b->mats.push_back(d);
        b->inputs.push_back(M.data);
    }
    // queue order: largest work first (m * wc^2 ~ sweep cost)
    b->order.resize(b->mats.size());
    std::iota(b->order.begin(), b->order.end(), 0);
    std::sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        const double cx = (double)b->mats[x].m * words64(b->mats[x].n) * words64(b->mats[x].n);
        const double cy = (double)b->mats[y].m * words64(b->mats[y].n) * words64(b->mats[y].n);
        return cx > cy;
    });
    b->work_words = work;
    b->z_words = z;
    const size_t nb = b->mats.size();
    mela_status st;
    if (flags & MELA_BATCH_CTX_SCRATCH) {
        // one batch at a time on this context: borrow its cached arenas
        b->ctx_scratch = true;
        void *desc = nullptr, *wk = nullptr, *za = nullptr, *rk = nullptr;
        const size_t desc_bytes =
            nb * (sizeof(BatchMat) + sizeof(void *) + sizeof(int) + MAX_GROUP * sizeof(int2) + sizeof(unsigned)) +
            64;
        if ((st = scratch_get(ctx, SLOT_BATCH_DESC, desc_bytes, &desc)) ||
            (st = scratch_get(ctx, SLOT_BATCH_WORK, work * 8, &wk)) ||
            (st = scratch_get(ctx, SLOT_BATCH_Z, z * 8, &za)) ||
            (st = scratch_get(ctx, SLOT_BATCH_RANK, nb * 4 + 4, &rk))) {
            delete b;
            return st;
        }
        char *dp = (char *)desc;
        b->d_mats = (BatchMat *)dp;
        dp += nb * sizeof(BatchMat);
        b->d_inputs = (const uint64_t **)dp;
        dp += nb * sizeof(void *);
        b->d_slots = (int2 *)dp;
        dp += nb * MAX_GROUP * sizeof(int2);
        b->d_order = (int *)dp;
        dp += nb * sizeof(int);
        b->d_gbar = (unsigned *)dp;
        dp += nb * sizeof(unsigned);
        b->d_counter = (int *)(((uintptr_t)dp + 15) & ~(uintptr_t)15);
        b->d_work = (uint64_t *)wk;
        b->d_z = (uint64_t *)za;
        b->d_ranks = (uint32_t *)rk;
    } else if ((st = dalloc(&b->d_mats, nb)) || (st = dalloc(&b->d_order, nb)) ||
               (st = dalloc(&b->d_counter, 2)) || (st = dalloc(&b->d_inputs, nb)) ||
               (st = dalloc(&b->d_slots, nb * MAX_GROUP)) || (st = dalloc(&b->d_gbar, nb)) ||
               (st = dalloc(&b->d_work, work)) || (st = dalloc(&b->d_z, z)) ||
               (st = dalloc(&b->d_ranks, nb))) {
        free_batch_device(b);
        delete b;
        return st;
    }
    hipStream_t s = ctx->stream;
    if (nb) {
        MELA_HIP(hipMemcpyAsync(b->d_order, b->order.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
        MELA_HIP(hipMemcpyAsync(b->d_inputs, b->inputs.data(), nb * sizeof(void *), hipMemcpyHostToDevice, s));
    }
    for (int tw : {K8_TW, 16, 32}) {
        if (tw == K8_TW && !g_k8_ok) continue;
        int occ = 0;
        MELA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_kernel(tw), batch_threads(tw),
                                                              batch_lds_bytes(tw)));
        b->occ_tw[tw_slot(tw)] = occ < 1 ? 1 : occ;
    }
    if ((st = batch_upload_plan(b)) != MELA_OK) {
        free_batch_device(b);
        delete b;
        return st;
    }
    *out = b;
    return MELA_OK;
}

namespace {

// Bodies of the setters, for callers that already hold the context lock.
mela_status set_groups_locked(mela_batch *b, const int *groups) {
    for (size_t i = 0; i < b->count; ++i)
        if (groups[i] < 1 || groups[i] > MAX_GROUP)
            return set_error(MELA_ERR_INVALID, "workgroups per matrix must be 1..8");
    for (size_t i = 0; i < b->count; ++i) b->mats[i].groups = (uint32_t)groups[i];
    return batch_upload_plan(b);
}

mela_status set_modes_locked(mela_batch *b, const int *modes) {
    for (size_t i = 0; i < b->count; ++i)
        if (modes[i] < 0 || modes[i] > 2)
            return set_error(MELA_ERR_INVALID, "batch modes are 0 (GE), 1 (M4RI k = 4) or 2 (M4RI k = 8)");
    for (size_t i = 0; i < b->count; ++i) b->mats[i].mode = (uint32_t)modes[i];
    return batch_upload_mats(b, b->ctx->stream);
}

mela_status autotune_locked(mela_batch *b, float *best_ms);

}  // namespace

mela_status mela_batch_set_groups(mela_batch *b, const int *groups) {
    if (!b || (b->count && !groups)) return set_error(MELA_ERR_INVALID, "null argument");
    CtxLock lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    return set_groups_locked(b, groups);
}

mela_status mela_batch_set_modes(mela_batch *b, const int *modes) {
    if (!b || !modes) return set_error(MELA_ERR_INVALID, "null argument");
    CtxLock lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    return set_modes_locked(b, modes);
}

mela_status mela_batch_set_tile(mela_batch *b, int tile_words) {
    if (!b) return set_error(MELA_ERR_INVALID, "null batch");
    if (tile_words != 8 && tile_words != 16 && tile_words != 32)
        return set_error(MELA_ERR_INVALID, "batch tile width must be 8, 16 or 32 words");
    if (tile_words == K8_TW && !g_k8_ok)
        return set_error(MELA_ERR_INVALID, "8-word batch tiles need more LDS per workgroup than this device grants");
    CtxLock lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    b->tw = tile_words;
    return batch_upload_plan(b);  // drops a graph captured with the old kernel and grid
}

mela_status mela_batch_get_plan(mela_batch *b, int *tile_words, int *modes, int *groups) {
    if (!b) return set_error(MELA_ERR_INVALID, "null batch");
    CtxLock lk(b->ctx->mu);
    if (tile_words) *tile_words = b->tw;
    for (size_t i = 0; i < b->count; ++i) {
        if (modes) modes[i] = (int)b->mats[i].mode;
        if (groups) groups[i] = (int)b->mats[i].groups;
    }
    return MELA_OK;
}

mela_status mela_batch_autotune(mela_batch *b, float *best_ms) {
    if (!b) return set_error(MELA_ERR_INVALID, "null batch");
    CtxLock lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    return autotune_locked(b, best_ms);
}

namespace {
mela_status autotune_locked(mela_batch *b, float *best_ms) {
    const size_t nb = b->mats.size();
    if (nb == 0) {
        if (best_ms) *best_ms = 0.f;
        return MELA_OK;
    }
    mela_ctx *ctx = b->ctx;
    hipStream_t s = ctx->stream;
    // per-matrix density (sparse parity-check matrices favour GE: few pivot
    // bits per row word; dense ones favour the 4-bit tables)
    std::vector<double> dens(nb);
    {
        unsigned long long *d_ones = nullptr;
        mela_status st = dalloc(&d_ones, nb);
        if (st != MELA_OK) return st;
        std::vector<unsigned long long> zero(nb, 0ull);
        MELA_HIP(hipMemcpyAsync(d_ones, zero.data(), nb * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_batch_popcount, dim3((unsigned)nb), dim3(256), 0, s, b->d_mats,
                           b->d_inputs, d_ones);
        MELA_HIP(hipGetLastError());
        MELA_HIP(hipMemcpyAsync(zero.data(), d_ones, nb * 8, hipMemcpyDeviceToHost, s));
        MELA_HIP(hipStreamSynchronize(s));
        (void)hipFree(d_ones);
        for (size_t i = 0; i < nb; ++i)
            dens[i] = (double)zero[i] / std::max(1.0, (double)b->mats[i].m * b->mats[i].n);
    }
    std::vector<std::vector<int>> cands;
    auto add = [&](auto pred) {
        std::vector<int> md(nb);
        for (size_t i = 0; i < nb; ++i) md[i] = pred(i) ? 1 : 0;
        cands.push_back(md);
    };
    add([&](size_t) { return false; });                               // all GE
    add([&](size_t) { return true; });                                // all M4RI
    add([&](size_t i) { return dens[i] > 0.05; });                     // dense -> M4RI
    add([&](size_t i) { return dens[i] > 0.05 && b->mats[i].n > 512; });
    add([&](size_t i) { return b->mats[i].n > 1024; });
    hipEvent_t e0 = nullptr, e1 = nullptr;
    struct EventPair {  // destroyed on every return, early ones included
        hipEvent_t &a, &b;
        ~EventPair() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } events{e0, e1};
    MELA_HIP(hipEventCreate(&e0));
    MELA_HIP(hipEventCreate(&e1));
    // time of the second of two runs (warm); a group wait time-out counts as
    // a failed candidate
    auto timed = [&](float *ms) -> mela_status {
        mela_status st;
        for (int rep = 0; rep < 2; ++rep) {
            MELA_HIP(hipEventRecord(e0, s));
            if ((st = batch_launch_raw(b, s)) != MELA_OK) return st;
            MELA_HIP(hipEventRecord(e1, s));
            MELA_HIP(hipEventSynchronize(e1));
            MELA_HIP(hipEventElapsedTime(ms, e0, e1));
        }
        int qerr[2] = {0, 0};
        MELA_HIP(hipMemcpy(qerr, b->d_counter, sizeof(qerr), hipMemcpyDeviceToHost));
        if (qerr[1]) *ms = 1e30f;
        return MELA_OK;
    };
    mela_status st;
    std::vector<int> ones(nb, 1);
    if ((st = set_groups_locked(b, ones.data())) != MELA_OK) return st;
    // search 1: tile width x per-matrix mode assignment.  The 16- and
    // 32-word tiles run M4RI with k = 4; the 8-word tiles also offer k = 8,
    // tried for every M4RI matrix of a candidate and for those of >= 512 rows
    // (a tile's 8-bit tables cost 32 LDS stores per thread, repaid by 8
    // fewer lookups per row word once a panel updates a few hundred rows).
    struct Plan {
        int tw;
        std::vector<int> md;
    };
    std::vector<Plan> plans;
    for (int tw : {16, 32})
        for (const auto &c : cands) plans.push_back({tw, c});
    for (const auto &c : cands)
        for (uint32_t rows : {0u, 512u}) {
            if (!g_k8_ok) break;
            std::vector<int> md = c;
            bool any = false;
            for (size_t i = 0; i < nb; ++i)
                if (md[i] && b->mats[i].m >= rows) {
                    md[i] = 2;
                    any = true;
                }
            if (any) plans.push_back({K8_TW, md});
        }
    float best = 1e30f;
    size_t bi = 0;
    for (size_t c = 0; c < plans.size(); ++c) {
        b->tw = plans[c].tw;
        if ((st = set_modes_locked(b, plans[c].md.data())) != MELA_OK) return st;
        float ms = 0.f;
        if ((st = timed(&ms)) != MELA_OK) return st;
        if (ms < best) {
            best = ms;
            bi = c;
        }
    }
    b->tw = plans[bi].tw;
    if ((st = set_modes_locked(b, plans[bi].md.data())) != MELA_OK) return st;
    b->tuned = (int)bi;
    // search 2: workgroups per matrix for the head of the queue (the largest
    // matrices, whose single-workgroup time bounds the launch)
    std::vector<int> gbest = ones;
    for (double frac : {0.005, 0.02, 0.06}) {
        for (int G : {2, 4, 8}) {
            std::vector<int> gr(nb, 1);
            const size_t top = std::max<size_t>(1, (size_t)(frac * (double)nb));
            for (size_t q = 0; q < top && q < nb; ++q) gr[b->order[q]] = G;
            if ((st = set_groups_locked(b, gr.data())) != MELA_OK) return st;
            float ms = 0.f;
            if ((st = timed(&ms)) != MELA_OK) return st;
            if (ms < best) {
                best = ms;
                gbest = gr;
            }
        }
    }
    // also drops a graph captured before tuning (old tile width and grid)
    if ((st = set_groups_locked(b, gbest.data())) != MELA_OK) return st;
    if (best_ms) *best_ms = best;
    return MELA_OK;
}
}  // namespace

mela_status mela_batch_run(mela_batch *b) {
    if (!b) return set_error(MELA_ERR_INVALID, "null batch");
    mela_ctx *ctx = b->ctx;
    CtxLock lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    if (!b->mats.empty()) {
        if (b->flags & MELA_BATCH_GRAPH) {
            if (!b->exec) {
                MELA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                mela_status st = batch_launch_raw(b, s);
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(s, &g);
                if (st != MELA_OK) return st;
                if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
                b->graph = g;
                MELA_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
            }
            MELA_HIP(hipGraphLaunch(b->exec, s));
        } else {
            mela_status st = batch_launch_raw(b, s);
            if (st != MELA_OK) return st;
        }
    }
    return MELA_OK;
}

mela_status mela_batch_results(mela_batch *b, uint32_t *ranks, mela_bitmat *nulls) {
    if (!b) return set_error(MELA_ERR_INVALID, "null batch");
    mela_ctx *ctx = b->ctx;
    {
        CtxLock lk(ctx->mu);
        MELA_HIP(hipStreamSynchronize(ctx->stream));
    }
    std::vector<uint32_t> r(b->mats.size());
    if (!r.empty()) {
        CtxLock lk(ctx->mu);
        int qerr[2] = {0, 0};
        MELA_HIP(hipMemcpyAsync(r.data(), b->d_ranks, r.size() * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, ctx->stream));
        MELA_HIP(hipMemcpyAsync(qerr, b->d_counter, sizeof(qerr), hipMemcpyDeviceToHost, ctx->stream));
        MELA_HIP(hipStreamSynchronize(ctx->stream));
        if (qerr[1])
            return set_error(MELA_ERR_DEVICE,
                             "a workgroup group of the batch timed out waiting for its partners "
                             "(the device was shared; rerun with one workgroup per matrix)");
    }
    for (size_t i = 0; i < b->count; ++i) {
        if (ranks) ranks[i] = r[i];
        if (nulls) {
            const BatchMat &d = b->mats[i];
            nulls[i].data = b->d_z + d.z_off;
            nulls[i].rows = d.n;
            nulls[i].cols = d.n - r[i];
            nulls[i].ld = d.ldz ? d.ldz : 1;
        }
    }
    return MELA_OK;
}

void mela_batch_destroy(mela_batch *b) {
    if (!b) return;
    if (b->ctx) {
        CtxLock lk(b->ctx->mu);
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        free_batch_device(b);
    } else {
        free_batch_device(b);
    }
    delete b;
}

}  // extern "C"
