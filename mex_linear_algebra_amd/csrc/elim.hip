// GF(2) elimination on the device: rank, reduced row echelon form, and the
// canonical null-space basis.
//
// The reference eliminates one pivot column at a time with a fork/join per
// pivot (rank_gf2.c:27-121, null_gf2.c:88-131) or per block of 8-10 pivots
// (rank_m4ri.c:94-279, null_m4ri.c:92-240).  Rank and RREF are unique, and the
// null basis built from the RREF (null_gf2.c:193-274) is the unique canonical
// one, so the device is free to choose a different elimination order; only
// the results must (and do) match.
//
// The device processes one 64-column panel (one packed word column g) at a
// time:
//   k_panel_select  one workgroup finds <= 64 active rows whose panel words
//                   span all active rows' panel words (wave Gauss-Jordan with
//                   ballot/readlane), removes them from the active list and
//                   computes the RREF basis of the panel plus, for each basis
//                   vector c, the set comb[c] of selected rows that XOR to it;
//   k_sweep         forms the panel's pivot rows of its tile from 4-bit tables
//                   of the selected rows (and writes them to Q), then clears
//                   the panel's pivot columns from every active row (and, for
//                   RREF, from every earlier pivot row) with one table XOR per
//                   4 pivot columns (M4RI) or one XOR per pivot row (GE),
//                   sweeping words g+1 .. wc-1; for RREF its tile-0 column
//                   also brings word g-1 of the earlier pivot rows up to date
//                   with panel g-1's pivot rows;
//   k_update_panelword (RREF) the same for the last panel's word.
#include <cstdio>
#include <mutex>

#include "elim_block.h"
#include "elim_device.h"

// Work memory of the allocation-free rank entry point
// (mela_rank_packed_noalloc): one copy per device, held until the process
// ends.  Host code never uses this name as an address; it takes the device
// address from hipGetSymbolAddress (noalloc_scratch_for_device).
constexpr size_t kNoAllocBytes = (size_t)32 << 20;
static __device__ __align__(256) uint64_t scratch[kNoAllocBytes / 8];

namespace mela {

// Column word of the null-basis gather (k_null_packed).
struct NullWord {
    uint64_t fm;     // free-column mask of the word
    uint64_t mv[6];  // its compress masks
    long fb;         // free columns before the word
    int ws;          // (entry j) the column word holding free column 64 j
    int pad;
};

namespace {

using dev::ffs64;
using dev::popc64;

constexpr int PS_THREADS = 1024;
// Sweep workgroups that record whether they left a nonzero active-row word
// past their panel: workgroup b of the sweep of panel g stores g + 1 into
// swept[b] (no reset needed between panels: older entries hold smaller tags).
constexpr int NSWEPT = 4096;

// One workgroup: selects the panel's pivot rows (elim_device.h) and records
// the panel in meta and pivcols.
__global__ void __launch_bounds__(PS_THREADS)
k_panel_select(const uint64_t *__restrict__ M, size_t ld, uint32_t *__restrict__ act,
               ElimMeta *__restrict__ meta, int *__restrict__ pivcols, const unsigned *swept,
               const uint64_t *__restrict__ panelw, int g, int rmax) {
    __shared__ dev::SelectLDS<PS_THREADS> L;
    __shared__ uint32_t actc[dev::ACT_HEAD + 64];
    // panel word g of active-list entry threadIdx.x as the last sweep left it
    // (it writes them when it ran, i.e. when the last selection found pivots):
    // loaded with the metadata instead of through act[] and the row
    const uint64_t pre = panelw[threadIdx.x];
    const uint64_t pm_old = meta->pm;
    const int r0_old = meta->r0;
    const bool have_pre = pm_old != 0;
    const int n_act = meta->n_act;
    const int r_before = meta->r;
    // no workgroup of the last sweep saw a nonzero active-row word past its
    // panel: no panel from here on has a pivot.  Checked after a panel with
    // fewer than 64 pivots (a full panel rarely ends the rank).
    bool finished = meta->finished != 0;
    const int ls = meta->last_swept;
    if (!finished && ls >= 0 && (meta->pm == 0 || meta->rho < 64)) {
        bool seen = false;
#pragma unroll
        for (int k = 0; k < NSWEPT / PS_THREADS; ++k)
            seen |= swept[threadIdx.x + k * PS_THREADS] == (unsigned)ls + 1u;
        finished = !__syncthreads_or(seen);
    }
    if (n_act == 0 || r_before >= rmax || finished) {
        if (threadIdx.x == 0) {
            meta->pm_prev = pm_old;
            meta->r0_prev = r0_old;
            meta->r0 = r_before;
            meta->rho = 0;
            meta->pm = 0;
            meta->finished = finished ? 1 : 0;
            if (meta->done_at > g) meta->done_at = g;
        }
        return;
    }
    auto rowfn = [&](int q) { return M + (size_t)q * ld; };
    const dev::PanelOut po = dev::wg_panel_select<PS_THREADS>(L, act, n_act, g, rowfn, meta->comb,
                                                              meta->sel_row, have_pre, pre, actc,
                                                              have_pre ? panelw : nullptr);
    const int tid = threadIdx.x;
    // slots past the selection keep a valid row id (the sweep loads every slot)
    if (tid < 64 && tid >= po.rho) meta->sel_row[tid] = 0;
    if (tid < 64 && ((po.pm >> tid) & 1ull))
        pivcols[r_before + popc64(po.pm & ((1ull << tid) - 1ull))] = g * 64 + tid;
    if (tid == 0) {
        meta->pm_prev = pm_old;
        meta->r0_prev = r0_old;
        meta->r0 = r_before;
        meta->rho = po.rho;
        meta->pm = po.pm;
        meta->r = r_before + po.rho;
        meta->n_act = po.n_act_new;
        if (po.pm) meta->last_swept = g;
    }
}

// A wave with a nonzero active-row word past panel g tags its workgroup's
// entry (plain stores of one value: no atomics, no contention; grids larger
// than NSWEPT share entries, which only merges tags of the same panel).
__device__ __forceinline__ void mark_rest_nonzero(unsigned *swept, uint64_t nz, int g) {
    const unsigned b = (blockIdx.y * gridDim.x + blockIdx.x) % NSWEPT;
    if (__ballot(nz != 0ull) != 0ull && (threadIdx.x & 63) == 0) swept[b] = (unsigned)g + 1u;
}

// Row pointer of unified row index q: active rows first, then (RREF) the
// earlier pivot rows.
__device__ __forceinline__ uint64_t *row_ptr(uint64_t *M, size_t ld, const uint32_t *act,
                                             uint64_t *Q, size_t ldq, int n_act, int q) {
    return q < n_act ? M + (size_t)act[q] * ld : Q + (size_t)(q - n_act) * ldq;
}

constexpr int UP_WAVES = 8;   // 512 threads
constexpr int UP_ROWS = 16;       // rows in flight per wave in the sweep
constexpr int SWEEP_TILE = 32;    // words per sweep workgroup (32 or 64)

// Row words that the next launch reads: written through the L2 (agent-scope
// relaxed stores, `sc1`).  The eight XCDs' L2s are not coherent with each
// other, so the end of a launch writes back every dirty L2 line for the
// workgroups of the next launch; stores that went through leave none, and
// the per-panel launches end sooner (3-4% over a 16384 x 16384 elimination).
__device__ __forceinline__ void store_through(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS-only workgroup barrier: row loads already in flight stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt((0xF) | (3 << 14) | (7 << 4) | (0 << 8));  // lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
}

// Sweep of panel g, pivot rows included: the workgroup's TILE-word tile starts
// at word g rounded down to 16 words (TILE 64: 128 KiB of tables, one
// workgroup per CU; TILE 32: two).
// Each thread owns TILE/32 (table t, word) pairs: it loads the selected rows
// 4t..4t+3 of its word and builds their 4-bit table t; pivot row c = XOR of
// the selected rows in comb[c] is then 16 lookups.  Row-stripe 0 of each
// tile writes the tile's pivot-row words to Q.  Each active (and for RREF
// earlier pivot) row then clears the panel's pivot columns from its words
// g+1 .. wc-1 (64 / TILE rows per wave instruction): M4RI XORs 16 entries of
// the pivot rows' 4-bit tables (the four pivot rows 4t..4t+3 a thread holds
// give table t; slot c = pivot bit c, zero where c is not a pivot), GE XORs
// the pivot row of every set bit of its panel word.  Word g of the earlier
// pivot rows is left to the next panel's sweep (or, for the last panel, to
// k_update_panelword) so that every tile reads an unmodified multiplier.
// MODE: SW_GE walks the pivot rows, SW_M4RI looks up the tables, SW_AUTO
// decides per row batch of a wave: a multiplier with more than
// AUTO_DENSE_BITS set bits on any of the wave's rows takes the 16 lookups,
// otherwise the walk (reading the pivot rows out of the tables, where entry
// 1 << b of table t is pivot row 4t + b).
enum { SW_GE = 0, SW_M4RI = 1, SW_AUTO = 2 };
constexpr int AUTO_DENSE_BITS = 8;
template <int TILE, int MODE>
__global__ void __launch_bounds__(UP_WAVES * 64, 64 / TILE)
k_sweep(uint64_t *__restrict__ M, size_t ld, const uint32_t *__restrict__ act,
        uint64_t *__restrict__ Q, size_t ldq, const ElimMeta *__restrict__ meta,
        unsigned *__restrict__ swept, uint64_t *__restrict__ panelw, int g, int wc, int rref,
        int mrows) {
    // T[entry][table][word]: the table index lands in the ds_read immediate
    // offset, the entry index is the only per-lookup address
    __shared__ uint64_t T[16][16][TILE];
    constexpr int NT = UP_WAVES * 64;
    constexpr int RPW = 64 / TILE;          // rows per wave instruction
    constexpr int R = UP_ROWS / RPW;        // row batch per lane
    constexpr int TPT = 16 * TILE / NT;     // (table, word) pairs per thread
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // RREF: word g-1 of the pivot rows found before panel g-1 still lacks
    // panel g-1's pivot rows (that sweep left word g-1 alone so that every
    // tile read an unmodified multiplier).  The tile-0 column applies them
    // here, where nothing else reads or writes that word; the loads go out
    // first and the update is done on the way out.
    __shared__ uint64_t PWd[64];
    const uint64_t pmp = rref && blockIdx.x == 0 && g > 0 ? meta->pm_prev : 0ull;
    const int rp0 = meta->r0_prev;
    const int dj0 = (int)blockIdx.y * NT + tid, dstride = (int)gridDim.y * NT;
    uint64_t dpw = 0, dx = 0;
    if (pmp) {
        if (tid < 64 && ((pmp >> tid) & 1ull))
            dpw = Q[(size_t)(rp0 + popc64(pmp & ((1ull << tid) - 1ull))) * ldq + (g - 1)];
        if (dj0 < rp0) dx = Q[(size_t)dj0 * ldq + (g - 1)];
    }
    auto finish_deferred = [&]() {  // every thread of the workgroup
        if (!pmp) return;
        if (tid < 64) PWd[tid] = dpw;
        __syncthreads();
        for (int j = dj0, it = 0; j < rp0; j += dstride, ++it) {
            const uint64_t xv = it == 0 ? dx : Q[(size_t)j * ldq + (g - 1)];
            uint64_t av = xv;
            for (uint64_t h = xv & pmp; h; h &= h - 1) av ^= PWd[ffs64(h)];
            store_through(&Q[(size_t)j * ldq + (g - 1)], av);
        }
    };
    const int wl = lane % TILE, sub = lane / TILE;
    const int per_wave = R * RPW;
    int q0 = (blockIdx.y * UP_WAVES + wave) * per_wave + sub;
    // (table, word) pairs of this thread: table e / TILE, word e % TILE
    auto pair_t = [&](int k) { return (tid + k * NT) / TILE; };
    // loads that need no metadata go out with it: the selected rows' ids
    // (every slot holds a valid row id; slots from rho on are discarded) and
    // the first row batch's active-list entries (every position below mrows
    // holds a row id)
    int srow[TPT][4];
#pragma unroll
    for (int k = 0; k < TPT; ++k)
#pragma unroll
        for (int b = 0; b < 4; ++b) srow[k][b] = meta->sel_row[4 * pair_t(k) + b];
    uint32_t pa[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int q = q0 + j * RPW;
        pa[j] = act[q < mrows ? q : mrows - 1];
    }
    const uint64_t pm = meta->pm;
    if (pm == 0) {
        finish_deferred();
        return;
    }
    const int n_act = meta->n_act, r0 = meta->r0, rho = meta->rho;
    const int total = n_act + (rref ? r0 : 0);
    // tiles start at word g rounded down to a 128-byte line (16 words), so
    // every row segment a tile reads or writes covers whole lines; words
    // below g+1 ride along unchanged
    const int g0 = g & ~15;
    const int w = g0 + blockIdx.x * TILE + wl;
    const bool win = w < wc;            // a word of the matrix
    const bool wok = win && w > g;      // a word the sweep updates
    const int wsafe = win ? w : g;      // loads of words past the end re-read word g
    const int stride = gridDim.y * UP_WAVES * per_wave;
    const bool rows_here = (int)blockIdx.y * UP_WAVES * per_wave < total;
    if (!rows_here && blockIdx.y != 0) {
        finish_deferred();
        return;
    }
    auto pair_w = [&](int k) { return g0 + blockIdx.x * TILE + (tid + k * NT) % TILE; };
    // selected rows' words first, then the first row batch, whose loads stay
    // in flight while the tables are built
    // the pivot-row combinations are loaded with the selected rows, so their
    // latency is spent while the first tables are built
    uint64_t cbv[TPT][4];
#pragma unroll
    for (int k = 0; k < TPT; ++k)
#pragma unroll
        for (int b = 0; b < 4; ++b) cbv[k][b] = meta->comb[4 * pair_t(k) + b];
    uint64_t sq[TPT][4];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int pw = pair_w(k) < wc ? pair_w(k) : g;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            sq[k][b] = M[(size_t)srow[k][b] * ld + pw];
        }
    }
    uint64_t *rp[R];
    uint64_t x[R], a[R];
    // first: the batch at q0, whose active-list entries are in pa (a row past
    // the end reads any valid row and is discarded)
    auto load_rows = [&](int qb, bool first) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int q = qb + j * RPW;
            if (first)
                rp[j] = q < n_act || q >= total ? M + (size_t)pa[j] * ld
                                                : Q + (size_t)(q - n_act) * ldq;
            else
                rp[j] = row_ptr(M, ld, act, Q, ldq, n_act, q < total ? q : 0);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) x[j] = rp[j][g];
#pragma unroll
        for (int j = 0; j < R; ++j) a[j] = rp[j][wsafe];
#pragma unroll
        for (int j = 0; j < R; ++j) x[j] = (qb + j * RPW < total && wok) ? x[j] & pm : 0ull;
    };
    if (rows_here) load_rows(q0, true);
    auto build = [&](const uint64_t (&v)[TPT][4]) {
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            const int t = pair_t(k), ww = (tid + k * NT) % TILE;
            uint64_t cur = 0;
            T[0][t][ww] = 0;
#pragma unroll
            for (int e = 1; e < 16; ++e) {
                cur ^= v[k][__builtin_ctz(e)];
                T[e ^ (e >> 1)][t][ww] = cur;  // Gray-code order
            }
        }
    };
#pragma unroll
    for (int k = 0; k < TPT; ++k)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (4 * pair_t(k) + b >= rho || pair_w(k) >= wc) sq[k][b] = 0ull;
    build(sq);  // tables of the selected rows
    lds_barrier();
    uint64_t pv[TPT][4];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int ww = (tid + k * NT) % TILE;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int c = 4 * pair_t(k) + b;
            const uint64_t cb = ((pm >> c) & 1ull) ? cbv[k][b] : 0ull;
            uint64_t v = 0;
#pragma unroll
            for (int u = 0; u < 16; ++u) v ^= T[(cb >> (4 * u)) & 15u][u][ww];
            pv[k][b] = v;
            if (blockIdx.y == 0 && pair_w(k) >= g && pair_w(k) < wc && ((pm >> c) & 1ull))
                store_through(&Q[(size_t)(r0 + popc64(pm & ((1ull << c) - 1ull))) * ldq + pair_w(k)], v);
        }
    }
    lds_barrier();
    if (!rows_here) {
        finish_deferred();
        return;
    }
    // GE: the pivot rows themselves, by pivot bit, over the tables' LDS
    uint64_t(*P)[TILE] = reinterpret_cast<uint64_t(*)[TILE]>(&T[0][0][0]);
    // pivot row c, word wl: its own slot (GE) or its table entry (AUTO)
    auto prow = [&](int c) -> uint64_t {
        return MODE == SW_GE ? P[c][wl] : T[1 << (c & 3)][c >> 2][wl];
    };
    if (MODE == SW_GE) {
#pragma unroll
        for (int k = 0; k < TPT; ++k)
#pragma unroll
            for (int b = 0; b < 4; ++b) P[4 * pair_t(k) + b][(tid + k * NT) % TILE] = pv[k][b];
    } else {
        build(pv);  // tables of the pivot rows
    }
    lds_barrier();
    uint64_t nz = 0;  // active rows' words past the panel, ORed
    // GE: the loop runs while any lane has rows (the wave's row halves end one
    // iteration apart at most), so the walk's wave reductions see every lane
    // active; a half past the end has zero multipliers
    for (; MODE != SW_M4RI ? __ballot(q0 < total) != 0ull : q0 < total; q0 += stride) {
        // rows whose multiplier is zero on the whole wave are skipped (a
        // wave-uniform branch); stores come after all lookups, so no row
        // waits for another row's store
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (__ballot(x[j] != 0ull) == 0ull) continue;
            const bool walk =
                MODE == SW_GE ||
                (MODE == SW_AUTO && __ballot(popc64(x[j]) > AUTO_DENSE_BITS) == 0ull);
            if (walk) {
                // one walk over the union of the wave's multipliers (its
                // TILE-lane rows), in scalar registers; a lane XORs pivot row
                // c only where its own multiplier has bit c (the sign-extended
                // bit is the mask), so every row still takes one XOR per set
                // multiplier bit, without per-lane loop control
                const uint32_t xl = (uint32_t)x[j], xh = (uint32_t)(x[j] >> 32);
                const uint32_t ul = dev::wave_or32(xl), uh = dev::wave_or32(xh);
                uint32_t al = (uint32_t)a[j], ah = (uint32_t)(a[j] >> 32);
                // two bits per step, so two pivot-row reads are in flight
                auto walk = [&](uint32_t h, uint32_t xs, int base) {
                    while (h) {
                        const int c1 = __builtin_ctz(h);
                        h &= h - 1;
                        const bool two = h != 0;
                        const int c2 = two ? __builtin_ctz(h) : c1;
                        if (two) h &= h - 1;
                        const uint64_t p1 = prow(base + c1), p2 = prow(base + c2);
                        const uint32_t m1 = (uint32_t)__builtin_amdgcn_sbfe((int)xs, c1, 1);
                        const uint32_t m2 =
                            two ? (uint32_t)__builtin_amdgcn_sbfe((int)xs, c2, 1) : 0u;
                        al ^= ((uint32_t)p1 & m1) ^ ((uint32_t)p2 & m2);
                        ah ^= ((uint32_t)(p1 >> 32) & m1) ^ ((uint32_t)(p2 >> 32) & m2);
                    }
                };
                walk(ul, xl, 0);
                walk(uh, xh, 32);
                a[j] = ((uint64_t)ah << 32) | al;
            } else {
#pragma unroll
                for (int t = 0; t < 16; ++t) a[j] ^= T[(x[j] >> (4 * t)) & 15u][t][wl];
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (x[j]) store_through(&rp[j][w], a[j]);
            if (q0 + j * RPW < n_act && wok) nz |= a[j];
        }
        // the next selection's panel words, in active-list order
        if (blockIdx.x == 0 && w == g + 1 && wok)
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int q = q0 + j * RPW;
                if (q < n_act) store_through(&panelw[q], a[j]);
            }
        if (q0 + stride < total) {
            load_rows(q0 + stride, false);
        } else if (MODE != SW_M4RI) {
#pragma unroll
            for (int j = 0; j < R; ++j) x[j] = 0ull;
        }
    }
    mark_rest_nonzero(swept, nz, g);
    finish_deferred();
}

// RREF only: word g of the earlier pivot rows (active rows' word g is dead),
// with the new pivot rows' word g staged in LDS by pivot bit.
__global__ void __launch_bounds__(256)
k_update_panelword(uint64_t *__restrict__ Q, size_t ldq, const ElimMeta *__restrict__ meta,
                   int g) {
    __shared__ uint64_t PW[64];
    const uint64_t pm = meta->pm;
    const int r0 = meta->r0;
    if (pm == 0) return;
    const int tid = threadIdx.x;
    if (tid < 64)
        PW[tid] = ((pm >> tid) & 1ull)
                      ? Q[(size_t)(r0 + popc64(pm & ((1ull << tid) - 1ull))) * ldq + g]
                      : 0ull;
    __syncthreads();
    const int q = blockIdx.x * 256 + tid;
    if (q >= r0) return;
    const uint64_t x = Q[(size_t)q * ldq + g];
    uint64_t a = x;
    for (uint64_t h = x & pm; h; h &= h - 1) a ^= PW[ffs64(h)];
    Q[(size_t)q * ldq + g] = a;
}

__global__ void k_elim_init(ElimMeta *meta, uint32_t *act, int m) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < m) act[i] = (uint32_t)i;
    if (i < 64) meta->sel_row[i] = 0;  // every slot holds a valid row id
    if (i == 0) {
        meta->n_act = m;
        meta->r = 0;
        meta->r0 = 0;
        meta->rho = 0;
        meta->pm = 0;
        meta->finished = 0;
        meta->last_swept = -1;
        meta->pm_prev = 0;
        meta->r0_prev = 0;
        meta->done_at = INT32_MAX;
    }
}

__global__ void __launch_bounds__(256)
k_copy_rows(const uint64_t *__restrict__ src, size_t lds, uint64_t *__restrict__ dst, size_t ldd,
            size_t rows, size_t wc) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= rows * wc) return;
    const size_t i = idx / wc, w = idx % wc;
    dst[i * ldd + w] = src[i * lds + w];
}

// Packed Z (n x f, row-major, ldz words): row p of column j is 1 if p is the
// j-th free column, bit free[j] of pivot row rowmap[p] if p is a pivot column
// (gf2null.m:53-55; null_gf2.c:223-273 computes the same entries by
// back-substitution).  One thread per word (p, jw), so the stores are
// coalesced: a free column's row is its unit vector; a pivot
// column's row gathers the free columns of its RREF row rowmap[p]: output
// word jw takes the compressed column words tab[jw].ws .. tab[jw+1].ws, each
// word's free bits packed by the six-step compress and shifted to their place
// (column word w holds free columns tab[w].fb ..).
__global__ void __launch_bounds__(256)
k_null_packed(const uint64_t *__restrict__ Q, size_t ldq, const int *__restrict__ rowmap,
              const NullWord *__restrict__ tab, size_t n, size_t wc, size_t f,
              uint64_t *__restrict__ Z, size_t ldz) {
    const size_t fw = (f + 63) / 64;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n * ldz; i += (size_t)gridDim.x * 256) {
        const size_t p = i / ldz, jw = i - p * ldz;
        uint64_t word = 0;
        if (jw < fw) {
            const int rm = rowmap[p];
            if (rm >= 0) {
                const uint64_t *row = Q + (size_t)rm * ldq;
                const long lo = 64 * (long)jw;
                const int wa = tab[jw].ws, wb = jw + 1 < fw ? tab[jw + 1].ws : (int)wc - 1;
                for (int w = wa; w <= wb; ++w) {
                    const long fb = tab[w].fb;
                    if (fb >= lo + 64) break;
                    const uint64_t c = dev::compress_apply(row[w], tab[w].fm, tab[w].mv);
                    const long sh = fb - lo;
                    word |= sh >= 0 ? c << sh : c >> (-sh);
                }
            } else {
                const size_t fi = (size_t)(-rm - 1);
                if (fi / 64 == jw) word = 1ull << (fi & 63);
            }
        }
        Z[i] = word;
    }
}

}  // namespace

namespace {
// Offsets inside the metadata buffer: ElimMeta, the pivot columns, the
// sweep tags, the next selection's panel words.
size_t meta_swept_off(size_t rmax) {
    return (sizeof(ElimMeta) + (rmax + 1) * sizeof(int) + 15) & ~(size_t)15;
}
size_t meta_panelw_off(size_t rmax) { return meta_swept_off(rmax) + NSWEPT * sizeof(unsigned); }
size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

// Shapes the blocked path takes: MELA_ELIM_BLOCKED always, AUTO from
// BLK_MIN_ROWS x BLK_MIN_COLS on (dense n x n: 0.23 / 0.43 / 1.68 ms blocked
// against 0.28 / 0.53 / 2.15 ms on the panel path at n = 1024 / 2048 / 8192).
constexpr size_t BLK_MIN_ROWS = 1024, BLK_MIN_COLS = 1024;
static bool blocked_candidate(int variant, size_t m, size_t n) {
    if (variant == MELA_ELIM_BLOCKED) return true;
    return variant == MELA_ELIM_AUTO && m >= BLK_MIN_ROWS && n >= BLK_MIN_COLS;
}

ElimLayout elim_layout(size_t m, size_t n) {
    ElimLayout L;
    const size_t wc = words64(n), rmax = m < n ? m : n;
    L.ld = (wc + 15) & ~(size_t)15;
    L.meta_bytes = round256(meta_panelw_off(rmax) + std::max<size_t>(m, PS_THREADS) * sizeof(uint64_t));
    L.act_off = L.meta_bytes;
    L.m_off = round256(L.act_off + m * sizeof(uint32_t));
    L.q_off = round256(L.m_off + m * L.ld * 8);
    L.total = round256(L.q_off + rmax * L.ld * 8);
    return L;
}

mela_status eliminate(mela_ctx *ctx, const uint64_t *A, size_t lda, size_t m, size_t n,
                      int variant, bool rref, bool want_pivots, ElimResult *res,
                      uint64_t **Q_out, size_t *ldq_out, const ElimBufs *fixed) {
    hipStream_t s = ctx->stream;
    res->rank = 0;
    res->pivots.clear();
    if (Q_out) *Q_out = nullptr;
    const size_t wc = words64(n);
    const size_t rmax = m < n ? m : n;
    // working rows (and the pivot rows Q) start on 128-byte lines, so the
    // sweep's line-aligned tiles stay aligned in every row
    const size_t ld = (wc + 15) & ~(size_t)15;
    if (ldq_out) *ldq_out = ld ? ld : 1;
    if (m == 0 || n == 0) return MELA_OK;
    if (m > (size_t)INT32_MAX || wc > (size_t)INT32_MAX)
        return set_error(MELA_ERR_INVALID, "matrix dimension exceeds 2^31-1");

    void *pM = nullptr, *pQ = nullptr, *pMeta = nullptr, *pAct = nullptr, *pBlk = nullptr;
    mela_status st;
    const size_t swept_off = meta_swept_off(rmax);
    const size_t panelw_off = meta_panelw_off(rmax);
    // the blocked path (elim_block.hip) for large matrices; never with the
    // caller's fixed buffers
    const bool blocked_shape = !fixed && blocked_candidate(variant, m, n);
    if (fixed) {
        pM = fixed->M;
        pQ = fixed->Q;
        pMeta = fixed->meta;
        pAct = fixed->act;
    } else {
        if ((st = scratch_get(ctx, SLOT_ELIM_M, m * ld * 8, &pM)) != MELA_OK) return st;
        if ((st = scratch_get(ctx, SLOT_ELIM_Q, rmax * ld * 8, &pQ)) != MELA_OK) return st;
        if ((st = scratch_get(ctx, SLOT_ELIM_META, panelw_off + std::max<size_t>(m, PS_THREADS) * sizeof(uint64_t),
                              &pMeta)) != MELA_OK)
            return st;
        if ((st = scratch_get(ctx, SLOT_ELIM_ACT, m * sizeof(uint32_t), &pAct)) != MELA_OK)
            return st;
        if (blocked_shape &&
            (st = scratch_get(ctx, SLOT_ELIM_BLK, block_scratch_bytes(m, n, rref), &pBlk)) != MELA_OK)
            return st;
    }
    uint64_t *Mw = (uint64_t *)pM, *Q = (uint64_t *)pQ;
    ElimMeta *meta = (ElimMeta *)pMeta;
    int *pivcols = (int *)((char *)pMeta + sizeof(ElimMeta));
    uint32_t *act = (uint32_t *)pAct;
    unsigned *swept = (unsigned *)((char *)pMeta + swept_off);
    uint64_t *panelw = (uint64_t *)((char *)pMeta + panelw_off);

    const size_t tot = m * wc;
    auto copy_rows = [&]() -> mela_status {
        hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, A, lda,
                           Mw, ld, m, wc);
        MELA_HIP(hipGetLastError());
        return MELA_OK;
    };
    if ((st = copy_rows()) != MELA_OK) return st;

    const int cu = ctx->cu_count > 0 ? ctx->cu_count : 256;
    // the panel sweep: GE and M4RI force theirs, every other variant the
    // per-batch choice
    const int sweep = variant == MELA_ELIM_GE ? SW_GE : variant == MELA_ELIM_M4RI ? SW_M4RI : SW_AUTO;
    auto init_state = [&]() -> mela_status {
        hipLaunchKernelGGL(k_elim_init, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, meta, act,
                           (int)m);
        // pivot rows are zero left of their panel (only their words >= g are
        // written); the RREF output and the null extraction read them whole
        if (rref) {
            mela_status zs = launch_fill_zero(s, Q, rmax * ld * sizeof(uint64_t));
            if (zs != MELA_OK) return zs;
        }
        return MELA_OK;
    };
    // panels g_begin .. g_end-1 (the first range also initialises the state);
    // `last`: for RREF, the final panel's word of the earlier pivot rows too
    auto enqueue_panels = [&](size_t g_begin, size_t g_end, bool last) -> mela_status {
        if (g_begin == 0) {
            mela_status zs = init_state();
            if (zs != MELA_OK) return zs;
            zs = launch_fill_zero(s, swept, NSWEPT * sizeof(unsigned));
            if (zs != MELA_OK) return zs;
        }
        for (size_t g = g_begin; g < g_end; ++g) {
            hipLaunchKernelGGL(k_panel_select, dim3(1), dim3(PS_THREADS), 0, s, Mw, ld, act, meta,
                               pivcols, swept, panelw, (int)g, (int)rmax);
            {
                // the sweep forms the panel's pivot rows of its tile itself;
                // tiles start at word g rounded down to 16, rows strided over
                // gridDim.y, one round of the resident workgroups (64 / TILE
                // per CU)
                const size_t rem = wc - (g & ~(size_t)15);
                const unsigned tx = (unsigned)((rem + SWEEP_TILE - 1) / SWEEP_TILE);
                const size_t rows_hint = m + (rref ? rmax : 0);
                unsigned ty = (unsigned)(((64 / SWEEP_TILE) * cu + tx - 1) / tx);
                const unsigned per = UP_WAVES * UP_ROWS;
                const unsigned ty_max = (unsigned)((rows_hint + per - 1) / per);
                if (ty > ty_max) ty = ty_max;
                if (ty < 1) ty = 1;
                if (sweep == SW_GE)
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sweep<SWEEP_TILE, SW_GE>), dim3(tx, ty),
                                       dim3(UP_WAVES * 64), 0, s, Mw, ld, act, Q, ld, meta, swept, panelw, (int)g,
                                       (int)wc, rref ? 1 : 0, (int)m);
                else if (sweep == SW_M4RI)
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sweep<SWEEP_TILE, SW_M4RI>), dim3(tx, ty),
                                       dim3(UP_WAVES * 64), 0, s, Mw, ld, act, Q, ld, meta, swept, panelw, (int)g,
                                       (int)wc, rref ? 1 : 0, (int)m);
                else
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sweep<SWEEP_TILE, SW_AUTO>), dim3(tx, ty),
                                       dim3(UP_WAVES * 64), 0, s, Mw, ld, act, Q, ld, meta, swept, panelw, (int)g,
                                       (int)wc, rref ? 1 : 0, (int)m);
            }
        }
        // RREF: word g of the earlier pivot rows is brought up to date by the
        // next panel's sweep; the last panel's word here
        if (rref && last)
            hipLaunchKernelGGL(k_update_panelword, dim3((unsigned)((rmax + 255) / 256)), dim3(256),
                               0, s, Q, ld, meta, (int)(g_end - 1));
        MELA_HIP(hipGetLastError());
        return MELA_OK;
    };
    const size_t nblk = (wc + BLK_WORDS - 1) / BLK_WORDS;
    // blocks b_begin .. b_end-1; `start`: the state's initialisation and
    // block 0's selection first (when b_begin == 0)
    auto enqueue_blk = [&](size_t b_begin, size_t b_end, bool start) -> mela_status {
        if (start && b_begin == 0) {
            mela_status zs = init_state();
            if (zs != MELA_OK) return zs;
        }
        return enqueue_blocks(ctx, Mw, ld, act, meta, pivcols, Q, pBlk, m, n, rref, b_begin, b_end, start);
    };
    ElimMeta hm;
    BlockShared hb;
    auto read_meta = [&](bool blk) -> mela_status {
        MELA_HIP(hipMemcpyAsync(&hm, meta, sizeof(ElimMeta), hipMemcpyDeviceToHost, s));
        if (blk) MELA_HIP(hipMemcpyAsync(&hb, pBlk, sizeof(BlockShared), hipMemcpyDeviceToHost, s));
        MELA_HIP(hipStreamSynchronize(s));
        return MELA_OK;
    };
    auto read_pivots = [&]() -> mela_status {
        const int r_dev = hm.r;
        res->rank = (size_t)r_dev;
        if (want_pivots && r_dev > 0) {
            res->pivots.resize((size_t)r_dev);
            MELA_HIP(hipMemcpyAsync(res->pivots.data(), pivcols, (size_t)r_dev * sizeof(int),
                                    hipMemcpyDeviceToHost, s));
            MELA_HIP(hipStreamSynchronize(s));
        }
        if (Q_out) *Q_out = Q;
        return MELA_OK;
    };
    if (fixed) {
        // caller's buffers: every panel enqueued directly, no graph
        if ((st = enqueue_panels(0, wc, true)) != MELA_OK) return st;
        if ((st = read_meta(false)) != MELA_OK) return st;
        ctx->elim_panel++;
        return read_pivots();
    }
    // replay a captured sequence for a shape seen before (the work buffers are
    // the context's cached scratch, so their addresses repeat; a buffer that
    // moved, e.g. the block state grown by an RREF of the same shape, makes a
    // new entry)
    ElimGraph *eg = nullptr;
    for (auto &e : ctx->elim_graphs)
        if (e.m == m && e.n == n && e.ld == ld && e.variant == variant && e.rref == (int)rref &&
            e.bufs[0] == Mw && e.bufs[1] == Q && e.bufs[2] == meta && e.bufs[3] == act && e.bufs[4] == pBlk)
            eg = &e;
    if (!eg) {
        // bounded cache: the oldest shape (and its executable graph) makes room
        constexpr size_t kMaxGraphs = 32;
        if (ctx->elim_graphs.size() >= kMaxGraphs) {
            ElimGraph &old = ctx->elim_graphs.front();
            if (old.exec) (void)hipGraphExecDestroy(old.exec);
            if (old.graph) (void)hipGraphDestroy(old.graph);
            ctx->elim_graphs.erase(ctx->elim_graphs.begin());
        }
        ElimGraph e;
        e.m = m, e.n = n, e.ld = ld, e.variant = variant, e.rref = rref ? 1 : 0;
        e.bufs[0] = Mw, e.bufs[1] = Q, e.bufs[2] = meta, e.bufs[3] = act, e.bufs[4] = pBlk;
        ctx->elim_graphs.push_back(e);
        eg = &ctx->elim_graphs.back();
    }
    const bool blocked = blocked_shape && !eg->blocked_failed;
    // work units: blocks or panels; a unit's selection records when it finds
    // the elimination complete
    const size_t units = blocked ? nblk : wc;
    auto enqueue_units = [&](size_t u0, size_t u1) -> mela_status {
        return blocked ? enqueue_blk(u0, u1, true) : enqueue_panels(u0, u1, u1 == wc);
    };
    auto done_unit = [&]() -> size_t {
        const int d = blocked ? hb.done_blk : hm.done_at;
        return d >= 0 && (size_t)d < units ? (size_t)d : units;
    };
    eg->seen++;
    if (!eg->exec && eg->seen >= 2 && wc <= 4096 && ctx->capture_ok) {
        // the captured sequence stops after the units the last run needed
        // (units past the one whose selection finds the elimination complete
        // are empty launches); a matrix that needs more continues below
        const size_t g_end = eg->panels_needed ? eg->panels_needed : units;
        MELA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        mela_status cst = enqueue_units(0, g_end);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (cst != MELA_OK) return cst;
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        eg->graph = graph;
        eg->g_end = g_end;
        MELA_HIP(hipGraphInstantiate(&eg->exec, graph, nullptr, nullptr, 0));
    }
    auto drop_graph = [&]() {
        if (eg->exec) (void)hipGraphExecDestroy(eg->exec);
        if (eg->graph) (void)hipGraphDestroy(eg->graph);
        eg->exec = nullptr;
        eg->graph = nullptr;
    };
    if (eg->exec) {
        MELA_HIP(hipGraphLaunch(eg->exec, s));
        if ((st = read_meta(blocked)) != MELA_OK) return st;
        if (eg->g_end < units && done_unit() >= eg->g_end && !(blocked && hb.redo)) {
            // not complete within the captured units: the rest uncaptured,
            // and the next call captures the whole sequence
            if ((st = enqueue_units(eg->g_end, units)) != MELA_OK) return st;
            if ((st = read_meta(blocked)) != MELA_OK) return st;
            drop_graph();
            eg->panels_needed = units;
            eg->seen = 1;
        }
    } else if (blocked && eg->seen == 1) {
        // a shape's first call: block 0's selection alone, so that a matrix
        // whose head is too sparse for the blocked path (it gives up there)
        // does not queue every block before it falls back
        if ((st = enqueue_blk(0, 0, true)) != MELA_OK) return st;
        if ((st = read_meta(true)) != MELA_OK) return st;
        if (!hb.redo) {
            if ((st = enqueue_blk(0, units, false)) != MELA_OK) return st;
            if ((st = read_meta(true)) != MELA_OK) return st;
        }
        eg->panels_needed = done_unit() < units ? done_unit() + 1 : units;
    } else {
        if ((st = enqueue_units(0, units)) != MELA_OK) return st;
        if ((st = read_meta(blocked)) != MELA_OK) return st;
        eg->panels_needed = done_unit() < units ? done_unit() + 1 : units;
    }
    if (blocked && hb.redo) {
        // a block's head missed pivots: the blocked result is void; this
        // shape takes the panel path from now on
        drop_graph();
        eg->blocked_failed = true;
        eg->seen = 0;
        eg->panels_needed = 0;
        eg->g_end = 0;
        if ((st = copy_rows()) != MELA_OK) return st;
        if ((st = enqueue_panels(0, wc, true)) != MELA_OK) return st;
        if ((st = read_meta(false)) != MELA_OK) return st;
        eg->panels_needed = (size_t)hm.done_at < wc ? (size_t)hm.done_at + 1 : wc;
        ctx->elim_fallback++;
    } else if (blocked) {
        ctx->elim_blocked++;
    } else {
        ctx->elim_panel++;
    }
    return read_pivots();
}

mela_status launch_null_packed(mela_ctx *ctx, const uint64_t *Q, size_t ldq, const int *rowmap,
                               const std::vector<int> &piv, size_t n, size_t f, uint64_t *Z,
                               size_t ldz) {
    if (n == 0 || ldz == 0) return MELA_OK;
    // per column word: free mask, its compress masks, free columns before it;
    // tab[j].ws = the column word holding free column 64 j
    const size_t wc = words64(n);
    std::vector<NullWord> tab(wc + 1);
    std::vector<uint64_t> pmask(wc, 0);
    for (int c : piv) pmask[(size_t)c >> 6] |= 1ull << (c & 63);
    long fb = 0;
    for (size_t w = 0; w < wc; ++w) {
        const size_t valid = n - 64 * w < 64 ? n - 64 * w : 64;
        const uint64_t fm = ~pmask[w] & (valid == 64 ? ~0ull : ((1ull << valid) - 1ull));
        NullWord &t = tab[w];
        t.fm = fm;
        dev::compress_masks(fm, t.mv);
        t.fb = fb;
        const long cnt = __builtin_popcountll(fm);
        const long j = (fb + 63) / 64;
        if (cnt && 64 * j < fb + cnt) tab[(size_t)j].ws = (int)w;
        fb += cnt;
    }
    void *p = nullptr;
    mela_status st = scratch_get(ctx, SLOT_NULL_TAB, tab.size() * sizeof(NullWord), &p);
    if (st != MELA_OK) return st;
    MELA_HIP(hipMemcpyAsync(p, tab.data(), tab.size() * sizeof(NullWord), hipMemcpyHostToDevice,
                            ctx->stream));
    const size_t total = n * ldz;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, (size_t)1 << 20);
    hipLaunchKernelGGL(k_null_packed, dim3(blocks), dim3(256), 0, ctx->stream, Q, ldq, rowmap,
                       (const NullWord *)p, n, wc, f, Z, ldz);
    MELA_HIP(hipGetLastError());
    // the table is host memory until the copy has run
    MELA_HIP(hipStreamSynchronize(ctx->stream));
    return MELA_OK;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
bool g_noalloc_done[kMaxDevices];
void *g_noalloc_ptr[kMaxDevices];
std::string g_noalloc_err[kMaxDevices];
// one allocation-free call at a time per device: they share its array
std::mutex g_noalloc_call_mu[kMaxDevices];
}  // namespace

void *noalloc_scratch_for_device(int device, std::string *err) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (device < 0 || device >= kMaxDevices) {
        if (err) *err = "device index beyond the allocation-free table";
        return nullptr;
    }
    if (!g_noalloc_done[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            g_noalloc_err[device] = std::string("hipGetSymbolAddress of the allocation-free "
                                                "scratch failed: ") + hipGetErrorString(e);
            ptr = nullptr;
        }
        g_noalloc_ptr[device] = ptr;
        g_noalloc_done[device] = true;
    }
    if (!g_noalloc_ptr[device] && err) *err = g_noalloc_err[device];
    return g_noalloc_ptr[device];
}

}  // namespace mela

using namespace mela;

extern "C" mela_status mela_rank_packed_noalloc(mela_ctx *ctx, const mela_bitmat *A, int variant,
                                                size_t *rank) {
    if (!ctx || !rank || !A) return set_error(MELA_ERR_INVALID, "null argument");
    if (A->rows && A->cols && !A->data) return set_error(MELA_ERR_INVALID, "null data in A");
    if (A->cols && A->ld < words64(A->cols)) return set_error(MELA_ERR_INVALID, "ld too small in A");
    if (variant < MELA_ELIM_GE || variant > MELA_ELIM_PANEL)
        return set_error(MELA_ERR_INVALID, "unknown elimination variant");
    *rank = 0;
    const size_t m = A->rows, n = A->cols;
    if (m == 0 || n == 0) return MELA_OK;
    if (m > (size_t)INT32_MAX || n > (size_t)INT32_MAX)
        return set_error(MELA_ERR_INVALID, "matrix dimension exceeds 2^31-1");
    const ElimLayout L = elim_layout(m, n);
    if (L.total > kNoAllocBytes) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "matrix %zu x %zu needs %zu bytes of work memory; the allocation-free rank "
                      "holds at most %zu (use mela_rank_packed)",
                      m, n, L.total, kNoAllocBytes);
        return set_error(MELA_ERR_INVALID, buf);
    }
    CtxLock lk(ctx->mu);
    MELA_HIP(hipSetDevice(ctx->device));
    char *base = (char *)ctx->noalloc_scratch;
    if (!base) return set_error(MELA_ERR_DEVICE, ctx->noalloc_err, "mela:device");
    std::lock_guard<std::mutex> lk2(g_noalloc_call_mu[ctx->device]);
    ElimBufs b;
    b.meta = base;
    b.act = (uint32_t *)(base + L.act_off);
    b.M = (uint64_t *)(base + L.m_off);
    b.Q = (uint64_t *)(base + L.q_off);
    ElimResult res;
    const mela_status st = eliminate(ctx, A->data, A->ld, m, n, variant, false, false, &res,
                                     nullptr, nullptr, &b);
    if (st != MELA_OK) return st;
    *rank = res.rank;
    return MELA_OK;
}
