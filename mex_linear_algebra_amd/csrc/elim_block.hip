// Blocked elimination of large dense matrices: rank, RREF and the pivot rows
// for the null basis, 512 columns (8 packed words) per block.
//
// The reference eliminates one pivot column at a time (rank_gf2.c:27-121,
// null_gf2.c:88-131) or blocks of k = 8-10 columns with a 2^k table applied
// to every row (rank_m4ri.c:120-275, null_m4ri.c:92-240).  Rank and RREF are
// unique, so the device may order the work differently; here a whole
// 512-column block is factored first and then applied to every remaining row
// in one matrix-core product, which moves 8x fewer bytes than eight panel
// sweeps and replaces their table lookups by MFMAs.
//
// Block b (words w0 = 8b .. w1):
//   block selection  one workgroup: the head of the active list (its first
//                   BLK_HEAD rows, block words only) is copied to LDS and
//                   eliminated panel by panel (the panel selection of
//                   elim_device.h, pivot rows by 4-bit tables, table sweeps
//                   of the other head rows; the next panel's selection runs
//                   on one wave while the others finish the sweep).  Every
//                   head row also carries 512 "slot" bits: selected row s
//                   gets bit s, and the sweeps XOR them along, so each pivot
//                   row ends up as a known combination T of the selected rows
//                   as they were when the block started.  The block's pivot
//                   rows (block words, RREF on the block's pivot columns) go
//                   to Q, T to a side buffer, and the active list without the
//                   selected rows to the block state (its head entries).
//   k_block_pform   the pivot rows' words past the block, P = T x S, with S
//                   the selected rows' words (4-bit tables over the slots),
//                   into Q; P~ (the block's pivot rows indexed by block
//                   column, zero rows at non-pivot columns) transposed into
//                   the matrix-core operand layout; the remaining rows' block
//                   words re-laid as the other operand; the new head of the
//                   active list written back.
//   trailing update every active row (and, for RREF, every earlier pivot
//                   row) ^= (its block words) x P~ on the matrix cores (fp4,
//                   K = 512).  Because P is in RREF on the block's pivot
//                   columns, a row's coefficients are its own bits there.
//                   An active row whose block words do not vanish had a
//                   pivot the head missed: the call then discards the
//                   blocked result and runs the panel path (elim.hip).
// k_block_step runs the trailing update of block b on all workgroups but
// one, which selects block b+1 meanwhile: the update tiles that hold the head
// rows' words of block b+1 run first, in a launch of their own, so the
// selection, one workgroup's latency-bound work, hides behind the product.
#include <climits>

#include "elim_device.h"
#include "elim_block.h"
#include "mfma_device.h"
#include "transpose_device.h"

namespace mela {
namespace {

using dev::ffs64;
using dev::popc64;

constexpr int BS_THREADS = 1024;
constexpr int HMAX = BLK_HEAD;
constexpr int BRW = 2 * BLK_WORDS;  // head row: block words, then 512 slot bits
constexpr int PF_TW = 4;            // words per k_block_pform workgroup (one per wave)

// Head rows are stored column-major (word w of head row h at H[w][h]) and the
// tables [table][word][entry], so that a wave whose lanes are 64 head rows (or
// 64 pivot rows) reads one word of each with consecutive addresses, and its
// table lookups (one table and word, 16 possible entries) hit distinct banks.
struct BlockLDS {
    uint64_t H[BRW][HMAX];       // head rows: block words | slot bits
    uint64_t Tb[16][BRW][16];    // 4-bit tables [table][word][entry]
    uint64_t Pn[BRW][64];        // the panel's pivot rows by pivot bit
    uint64_t xs[HMAX];           // sweep multipliers
    dev::SelectLDS<BS_THREADS> S;
    uint64_t comb[64];
    uint64_t pmb[BLK_WORDS];
    int selrow[64];
    int kc[64];                  // k-th pivot bit of the panel
    uint32_t hl[HMAX];           // candidate list (head positions not yet selected)
    uint32_t hrow[HMAX];         // row id of each head position
    int16_t hslot[HMAX];         // selection slot of each head position, -1
    int16_t hpiv[HMAX];          // block column of its pivot, -1
    int16_t hidx[HMAX];          // its pivot index in the block, -1
    uint32_t movers[BLK_COLS];
    int wcnt[BS_THREADS / 64];
    int fsflag[64];              // wave selection: tail position taken
    uint32_t fsmv[64];           // wave selection: entries that move
    int fs_ok;                   // wave selection succeeded
};

// Word g of head row q, the row accessor the panel selection takes.
struct HeadRow {
    const uint64_t (*H)[HMAX];
    int q;
    __device__ uint64_t operator[](int g) const { return H[g][q]; }
};

// Panel selection by one wave (the fast path of dev::wg_panel_select): a
// tracked Gauss-Jordan over the first 64 candidates' word g, then single
// insertions of the next candidates (up to position dev::ACT_HEAD) until 64
// pivots.  On success the 64 selected entries leave hl (unselected entries of
// its last 64 positions fill their places), selrow[k] is the head position of
// selection slot k and comb[c] the slots XORed into pivot row c; without 64
// pivots it changes nothing and returns false.  All 64 lanes of one wave call
// it.
__device__ bool wave_fast_select(BlockLDS &L, int n, int g) {
    const int lane = threadIdx.x & 63;
    if (n < 64) return false;
    const uint64_t x0 = L.H[g][L.hl[lane]];
    if (popc64(__ballot(x0 != 0ull)) < dev::FAST_MIN_PIVOTS) return false;
    uint64_t v = x0, comb = 1ull << lane;
    bool piv;
    int pbit;
    dev::wave_gj<true>(v, comb, true, piv, pbit, lane);
    uint64_t pivl = __ballot(piv);
    int slotpos = lane;
    const int lim = popc64(pivl) < dev::FAST_MIN_PIVOTS ? 0 : min(n, dev::ACT_HEAD);
    for (int pos = 64; pos < lim && pivl != ~0ull; ++pos) {
        const uint64_t x = L.H[g][L.hl[pos]];  // the same on every lane
        if (x == 0) continue;
        const bool use = piv && ((x >> pbit) & 1ull);
        const uint64_t red = dev::wave_xor64(use ? v : 0ull);
        const uint64_t cred = dev::wave_xor64(use ? comb : 0ull);
        const uint64_t z = x ^ red;  // x reduced by the RREF basis
        if (z == 0) continue;
        const int sl = ffs64(~pivl);  // a free slot
        const uint64_t cz = cred ^ (1ull << sl);
        const int pb = ffs64(z);
        if (piv && ((v >> pb) & 1ull)) {
            v ^= z;
            comb ^= cz;
        }
        if (lane == sl) {
            v = z;
            comb = cz;
            piv = true;
            pbit = pb;
            slotpos = pos;
        }
        pivl |= 1ull << sl;
    }
    if (pivl != ~0ull) return false;
    // 64 pivots, slot = lane
    L.comb[pbit] = comb;
    const int n_new = n - 64;
    L.selrow[lane] = (int)L.hl[slotpos];
    L.fsflag[lane] = 0;
    __threadfence_block();
    if (slotpos >= n_new) L.fsflag[slotpos - n_new] = 1;
    __threadfence_block();
    const bool is_hole = slotpos < n_new;
    const bool is_mover = L.fsflag[lane] == 0;
    const uint64_t hb = __ballot(is_hole), mb = __ballot(is_mover);
    const uint64_t below = (1ull << lane) - 1ull;
    if (is_mover) L.fsmv[popc64(mb & below)] = L.hl[n_new + lane];
    __threadfence_block();
    if (is_hole) L.hl[slotpos] = L.fsmv[popc64(hb & below)];
    return true;
}

// Exclusive prefix count of `flag` over the workgroup (every thread calls).
__device__ __forceinline__ int wg_prefix(bool flag, int *wcnt, int *total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t bal = __ballot(flag);
    if (lane == 0) wcnt[wave] = popc64(bal);
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < BS_THREADS / 64; ++q) {
        const int c = wcnt[q];
        if (q < wave) off += c;
        tot += c;
    }
    __syncthreads();
    if (total) *total = tot;
    return off + popc64(bal & ((1ull << lane) - 1ull));
}

// Workgroup barrier that waits only for the LDS: global loads issued for the
// next tile stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt((0xF) | (3 << 14) | (7 << 4) | (0 << 8));  // lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
}

// Pivot index of block column bc (pivots of the block with smaller columns).
__device__ __forceinline__ int block_pivot_index(const uint64_t *pmb, int bc) {
    const int q = bc >> 6, c = bc & 63;
    int i = popc64(pmb[q] & ((1ull << c) - 1ull));
    for (int qq = 0; qq < q; ++qq) i += popc64(pmb[qq]);
    return i;
}

__global__ void k_block_init(BlockShared *sh, BlockState *bst) {
    if (threadIdx.x == 0) {
        sh->redo = 0;
        sh->alive = 0;
        sh->done_blk = INT_MAX;
        bst[0].active = 0;
        bst[1].active = 0;
    }
}

// Selection of block b into bs (all BS_THREADS threads of the workgroup).
__device__ __forceinline__ void block_select(BlockLDS &L, const uint64_t *__restrict__ M, size_t ld,
                             const uint32_t *__restrict__ act, ElimMeta *__restrict__ meta,
                             int *__restrict__ pivcols, BlockShared *__restrict__ sh,
                             BlockState *__restrict__ bs, uint64_t *__restrict__ Tout,
                             uint64_t *__restrict__ Q, size_t ldq, int b, int wc, int rmax) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int w0 = b * BLK_WORDS;
    const int nbw = min(BLK_WORDS, wc - w0);
    const int n_act = meta->n_act, r0 = meta->r;
    // the update of block b' tags alive with b'+1 when an active row kept a
    // nonzero word past block b'.  The update of block b-1 may still be
    // running; that of block b-2 is done, so a tag below b-1 means that every
    // active row was zero after block b-2.
    const bool finished = n_act == 0 || r0 >= rmax || (b >= 2 && sh->alive < b - 1);
    if (sh->redo || finished) {
        if (tid == 0) {
            bs->rb = 0;
            bs->r0 = r0;
            bs->n_new = n_act;
            bs->hn = 0;
            bs->active = 0;
            if (finished && !sh->redo && sh->done_blk > b) sh->done_blk = b;
        }
        return;
    }
    const int Hn = min(n_act, HMAX);
    for (int h = tid; h < Hn; h += BS_THREADS) {
        L.hrow[h] = act[h];
        L.hl[h] = (uint32_t)h;
        L.hslot[h] = -1;
        L.hpiv[h] = -1;
    }
    if (tid < BLK_WORDS) L.pmb[tid] = 0;
    __syncthreads();
    {
        // block words of the head: every load of a thread in flight at once
        constexpr int PER = (HMAX * BLK_WORDS + BS_THREADS - 1) / BS_THREADS;
        uint64_t v[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int idx = tid + k * BS_THREADS, h = idx / BLK_WORDS, w = idx - h * BLK_WORDS;
            v[k] = h < Hn && w < nbw ? M[(size_t)L.hrow[h] * ld + w0 + w] : 0ull;
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int idx = tid + k * BS_THREADS, h = idx / BLK_WORDS, w = idx - h * BLK_WORDS;
            if (h < Hn) L.H[w][h] = v[k];
        }
        for (int idx = tid; idx < Hn * BLK_WORDS; idx += BS_THREADS) {
            const int w = idx / Hn, h = idx - w * Hn;
            L.H[BLK_WORDS + w][h] = 0ull;
        }
    }
    __syncthreads();

    int n_cand = Hn, rb = 0;
    auto rowfn = [&](int q) { return HeadRow{L.H, q}; };
    // a sparse matrix shows in its head: when fewer than a quarter of the
    // head rows have a nonzero word in block 0 while other rows remain, the
    // head would miss pivots; give up at once (the call takes the panel path).
    // Dense matrices of any rank have about half of their rows nonzero there
    // or more; column-weight-3 matrices a few percent.
    if (b == 0 && Hn < n_act) {
        bool nz = false;
        if (tid < Hn)
#pragma unroll
            for (int w = 0; w < BLK_WORDS; ++w) nz |= L.H[w][tid] != 0ull;
        int cnt = 0;
        (void)wg_prefix(nz, L.wcnt, &cnt);
        if (4 * cnt < Hn) {
            if (tid == 0) {
                sh->redo = 1;
                bs->active = 0;
            }
            return;
        }
    }
    // selection of panel 0 by the whole workgroup; each later panel's
    // selection overlaps the previous panel's sweep (below)
    uint64_t pm = 0;
    int rho = 0;
    if (nbw > 0 && n_cand > 0) {
        const dev::PanelOut po =
            dev::wg_panel_select<BS_THREADS>(L.S, L.hl, n_cand, 0, rowfn, L.comb, L.selrow);
        pm = po.pm;
        rho = po.rho;
        n_cand = po.n_act_new;
    }
    for (int p = 0; p < nbw; ++p) {
        // words in use: block words p .. nbw-1, then the slot words holding
        // slots 0 .. rb + rho - 1 (the others are zero in every head row)
        const int twn = (rb + rho + 63) >> 6;
        const int nw1 = nbw - p, nwd = nw1 + twn;
        auto word_of = [&](int wi) { return wi < nw1 ? p + wi : BLK_WORDS + (wi - nw1); };
        // x / nwd for 0 <= x < 256 (nwd <= 16) by a multiply and shift: the
        // rounding error x * (ceil(2^16 / nwd) - 2^16 / nwd) < 256 stays
        // below the 2^16 / nwd gap between multiples
        const uint32_t nwd_inv = (65536u + (uint32_t)nwd - 1u) / (uint32_t)nwd;
        auto div_nwd = [&](int x) { return (int)(((uint32_t)x * nwd_inv) >> 16); };
        // sweep of head rows h = i0, i0 + step, ... over word w with the
        // panel's tables (multipliers in xs)
        auto sweep_word = [&](int w, int i0, int step) {
            for (int h = i0; h < Hn; h += step) {
                const uint64_t x = L.xs[h];
                if (x == 0) continue;
                uint64_t a = L.H[w][h];
#pragma unroll
                for (int u = 0; u < 16; ++u) a ^= L.Tb[u][w][(x >> (4 * u)) & 15u];
                L.H[w][h] = a;
            }
        };
        const bool next = p + 1 < nbw;  // word p+1 is wi == 1
        if (rho > 0) {
            // slots rb .. rb + rho - 1 in selection order; each selected
            // row's slot bits get its own slot
            if (tid < rho) {
                const int h = L.selrow[tid], s = rb + tid;
                L.hslot[h] = (int16_t)s;
                L.H[BLK_WORDS + (s >> 6)][h] ^= 1ull << (s & 63);
                bs->slot_row[s] = L.hrow[h];
            }
            if (tid == 0) L.pmb[p] = pm;
            __syncthreads();
            // 4-bit tables of the selected rows (table t: selection slots 4t
            // .. 4t+3), one entry per thread (lanes = entries: consecutive
            // stores); the four rows' words are read unconditionally, so
            // their loads issue together
            for (int idx = tid; idx < 16 * 16 * nwd; idx += BS_THREADS) {
                const int e = idx & 15, tw = idx >> 4, t = div_nwd(tw), w = word_of(tw - t * nwd);
                uint64_t a[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int k = 4 * t + q;
                    a[q] = L.H[w][L.selrow[k < rho ? k : 0]];
                }
                uint64_t v = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) v ^= (((e >> q) & 1) && 4 * t + q < rho) ? a[q] : 0ull;
                L.Tb[t][w][e] = v;
            }
            __syncthreads();
            // the panel's pivot rows: pivot bit c = XOR of the selected rows
            // in comb[c] (lanes = pivot bits, one word per wave)
            for (int idx = tid; idx < 64 * nwd; idx += BS_THREADS) {
                const int w = word_of(idx >> 6), c = idx & 63;
                uint64_t v = 0;
                if ((pm >> c) & 1ull) {
                    const uint64_t cb = L.comb[c];
#pragma unroll
                    for (int u = 0; u < 16; ++u) v ^= L.Tb[u][w][(cb >> (4 * u)) & 15u];
                    if (idx < 64) L.kc[popc64(pm & ((1ull << c) - 1ull))] = c;
                }
                L.Pn[w][c] = v;
            }
            __syncthreads();
            // sweep tables of the pivot rows (table t: pivot bits 4t .. 4t+3)
            // and the multipliers of every head entry but this panel's
            // selected ones
            for (int idx = tid; idx < 16 * 16 * nwd; idx += BS_THREADS) {
                const int e = idx & 15, tw = idx >> 4, t = div_nwd(tw), w = word_of(tw - t * nwd);
                uint64_t a[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) a[q] = L.Pn[w][4 * t + q];
                uint64_t v = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) v ^= ((e >> q) & 1) ? a[q] : 0ull;
                L.Tb[t][w][e] = v;
            }
            for (int h = tid; h < Hn; h += BS_THREADS) {
                const int s = L.hslot[h];
                const bool now = s >= rb && s < rb + rho;
                L.xs[h] = now ? 0ull : (L.H[p][h] & pm);
            }
            __syncthreads();
            // this panel's selected entries become its pivot rows (entry
            // selrow[k] takes the k-th pivot bit's row), and word p+1 of
            // every other entry is swept first: the next selection reads it
            if (tid < rho * nwd) {
                const int wi = tid / rho, k = tid - wi * rho;
                const int w = word_of(wi);
                const int c = L.kc[k];
                const int h = L.selrow[k];
                L.H[w][h] = L.Pn[w][c];
                if (wi == 0) L.hpiv[h] = (int16_t)(64 * p + c);
            }
            if (next) sweep_word(p + 1, tid, BS_THREADS);
            rb += rho;
            __syncthreads();
        }
        // Block 0's first word with fewer than 64 pivots and rows past the
        // head: 512 rows spread over the rest of the active list, reduced by
        // the panel's pivot rows (Pn, RREF in word 0); a nonzero word is a
        // pivot the head missed, found here after one panel (the check after
        // the last panel covers the other words)
        if (b == 0 && p == 0 && rho < 64 && Hn < n_act) {
            const int pos = Hn + (int)((long long)tid * (n_act - Hn) / BS_THREADS);
            const uint64_t x = M[(size_t)act[pos] * ld + w0];
            uint64_t r = x;
            for (uint64_t mb = x & pm; mb; mb &= mb - 1) r ^= L.Pn[0][ffs64(mb)];
            if (__syncthreads_or(r != 0ull)) {
                if (tid == 0) {
                    sh->redo = 1;
                    bs->active = 0;
                }
                return;
            }
        }
        // the rest of the sweep (word p and the words past p+1) on waves
        // 1..15 while wave 0 selects panel p+1 from the candidates' word p+1
        // (the fast path; the whole workgroup takes the general path when
        // it declines)
        const int rho_now = rho;
        if (next && n_cand > 0) {
            if (wave == 0) {
                const bool ok = wave_fast_select(L, n_cand, p + 1);
                if (lane == 0) L.fs_ok = ok ? 1 : 0;
            } else if (rho_now > 0) {
                for (int idx = tid - 64; idx < HMAX * (nwd - 1); idx += BS_THREADS - 64) {
                    const int j = idx / HMAX, h = idx - j * HMAX;
                    const int wi = j == 0 ? 0 : j + 1;
                    if (h < Hn) sweep_word(word_of(wi), h, HMAX);
                }
            }
            __syncthreads();
            if (L.fs_ok) {
                pm = ~0ull;
                rho = 64;
                n_cand -= 64;
            } else {
                const dev::PanelOut po = dev::wg_panel_select<BS_THREADS>(L.S, L.hl, n_cand, p + 1, rowfn,
                                                                          L.comb, L.selrow);
                pm = po.pm;
                rho = po.rho;
                n_cand = po.n_act_new;
            }
        } else {
            if (rho_now > 0) {
                // no later selection in this block: only the earlier pivot
                // rows (which must end in RREF) still need the sweep
                for (int idx = tid; idx < HMAX * nwd; idx += BS_THREADS) {
                    const int wi = idx / HMAX, h = idx - wi * HMAX;
                    if (wi == 1 && next) continue;
                    if (h < Hn && L.hslot[h] >= 0) sweep_word(word_of(wi), h, HMAX);
                }
                __syncthreads();
            }
            pm = 0;
            rho = 0;
        }
    }

    // Block 0 with fewer pivots than columns and rows past the head: reduce 64
    // rows spread over the rest of the active list by the head's pivot rows
    // (in RREF, so row x reduces to x ^ XOR of P_c over the pivot columns c
    // where x is set).  A nonzero remainder is a pivot the head missed, which
    // the update would only report after it ran; give up now (the call takes
    // the panel path).  Dense blocks of full rank skip this.
    if (b == 0 && Hn < n_act && rb < 64 * nbw) {
        for (int h = tid; h < Hn; h += BS_THREADS)
            if (L.hpiv[h] >= 0) L.movers[L.hpiv[h]] = (uint32_t)h;  // pivot column -> head position
        __syncthreads();
        bool miss = false;
        if (tid < 64 * BLK_WORDS) {
            const int smp = tid / BLK_WORDS, j = tid % BLK_WORDS;
            const int pos = Hn + (int)((long long)smp * (n_act - Hn) / 64);
            const size_t row = act[pos];
            uint64_t xw[BLK_WORDS];
#pragma unroll
            for (int q = 0; q < BLK_WORDS; ++q) xw[q] = q < nbw ? M[row * ld + w0 + q] : 0ull;
            uint64_t r = 0;
#pragma unroll
            for (int q = 0; q < BLK_WORDS; ++q)
                if (q == j) r = xw[q];
            if (j < nbw) {
                for (int q = 0; q < nbw; ++q)
                    for (uint64_t mb = xw[q] & L.pmb[q]; mb; mb &= mb - 1)
                        r ^= L.H[j][L.movers[64 * q + ffs64(mb)]];
                miss = r != 0ull;
            }
        }
        if (__syncthreads_or(miss)) {
            if (tid == 0) {
                sh->redo = 1;
                bs->active = 0;
            }
            return;
        }
    }

    // pivot rows in block-column order: block words to Q, slot bits to Tout
    for (int h = tid; h < Hn; h += BS_THREADS) {
        const int bc = L.hpiv[h];
        L.hidx[h] = (int16_t)(bc < 0 ? -1 : block_pivot_index(L.pmb, bc));
    }
    __syncthreads();
    for (int idx = tid; idx < Hn * BRW; idx += BS_THREADS) {
        const int h = idx / BRW, w = idx - h * BRW;
        const int i = L.hidx[h];
        if (i < 0) continue;
        const int bc = L.hpiv[h];
        if (w < BLK_WORDS) {
            if (w < nbw) Q[(size_t)(r0 + i) * ldq + w0 + w] = L.H[w][h];
        } else {
            Tout[(size_t)i * BLK_WORDS + (w - BLK_WORDS)] = L.H[w][h];
        }
        if (w == 0) pivcols[r0 + i] = 64 * w0 + bc;
    }
    // the selected rows leave the active list: the unselected entries of
    // positions n_new .. n_act-1 (in order) fill the selected positions below
    // n_new (in order); positions >= Hn were never selected.  The new head
    // (entries 0 .. Hn-1) goes to bs->headmap; k_block_pform writes it back.
    const int n_new = n_act - rb;
    {
        const int pos = n_new + tid;
        bool mv = false;
        uint32_t val = 0;
        if (tid < rb) {
            mv = !(pos < Hn && L.hslot[pos] >= 0);
            if (mv) val = pos < Hn ? L.hrow[pos] : act[pos];
        }
        const int k = wg_prefix(mv, L.wcnt, nullptr);
        if (mv) L.movers[k] = val;
    }
    __syncthreads();
    {
        const bool hole = tid < Hn && tid < n_new && L.hslot[tid] >= 0;
        const int k = wg_prefix(hole, L.wcnt, nullptr);
        if (tid < Hn) bs->headmap[tid] = hole ? L.movers[k] : L.hrow[tid];
    }
    if (tid < BLK_WORDS) bs->pmb[tid] = L.pmb[tid];
    if (tid == 0) {
        bs->rb = rb;
        bs->r0 = r0;
        bs->n_new = n_new;
        bs->hn = Hn;
        bs->active = 1;
        meta->r = r0 + rb;
        meta->n_act = n_new;
    }
}

// Selection of block 0 (the later blocks' selections run inside k_block_step).
__global__ void __launch_bounds__(BS_THREADS)
k_block_select(const uint64_t *__restrict__ M, size_t ld, const uint32_t *__restrict__ act,
               ElimMeta *__restrict__ meta, int *__restrict__ pivcols, BlockShared *__restrict__ sh,
               BlockState *__restrict__ bst, uint64_t *__restrict__ Tout, uint64_t *__restrict__ Q,
               size_t ldq, int wc, int rmax) {
    __shared__ BlockLDS L;
    block_select(L, M, ld, act, meta, pivcols, sh, bst, Tout, Q, ldq, 0, wc, rmax);
}

// Workgroups 0 .. 8*ntw-1: pivot rows of the block's panel q = blockIdx % 8
// over words wa .. wa+3 (wa = w0 + 4 (blockIdx / 8)); one wave per word, lane
// = pivot bit.  Words inside the block come from Q (the selection wrote
// them), words past it are P = T x S with 4-bit tables of the selected rows'
// words (table t: slots 4t .. 4t+3, laid out [table][entry][word] so that a
// wave's lookups, one table and word with 16 possible entries, hit distinct
// banks).  Each wave then transposes its 64 x 64 bit block: lane j gets
// column 64 (w - w0) + j's bits at block columns 64q .. 64q+63, which is that
// column's 8 bytes of the P~ operand.
// Workgroups from 8*ntw on: the other operand, the block words of the rows
// the update covers (active rows in list order, then for RREF the pivot rows
// found before the block), 32 bytes per (row, stage), zero rows up to a
// multiple of 256; the first of them also writes the block's new head of the
// active list into act (the others read it from bs->headmap meanwhile).
__global__ void __launch_bounds__(256)
k_block_pform(const uint64_t *__restrict__ M, size_t ld, uint32_t *__restrict__ act,
              const BlockShared *__restrict__ sh, const BlockState *__restrict__ bs,
              const uint64_t *__restrict__ Tin, uint64_t *__restrict__ Q, size_t ldq,
              uint8_t *__restrict__ Xa, uint8_t *__restrict__ Bt, int b, int wc, int ntw, int rref,
              int xa_blocks) {
    __shared__ uint64_t Tbl[BLK_COLS / 4][16][PF_TW];
    if (sh->redo || !bs->active) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int w0 = b * BLK_WORDS, w1 = min(w0 + BLK_WORDS, wc);
    const int r0 = bs->r0;
    if ((int)blockIdx.x >= 8 * ntw) {
        const int n_new = bs->n_new, hn = bs->hn;
        const int rows = n_new + (rref ? r0 : 0);
        const int rows_pad = (rows + 255) & ~255;
        if ((int)blockIdx.x == 8 * ntw)
            for (int i = tid; i < hn; i += 256) act[i] = bs->headmap[i];
        for (int idx = ((int)blockIdx.x - 8 * ntw) * 256 + tid; idx < 2 * rows_pad;
             idx += xa_blocks * 256) {
            const int i = idx >> 1, s = idx & 1;
            uint64_t v[4] = {0ull, 0ull, 0ull, 0ull};
            if (i < rows) {
                const uint32_t a = i < hn ? bs->headmap[i] : act[i];
                const uint64_t *src = i < n_new ? M + (size_t)a * ld : Q + (size_t)(i - n_new) * ldq;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int w = w0 + 4 * s + q;
                    v[q] = w < w1 ? src[w] : 0ull;
                }
            }
            uint4 *p = reinterpret_cast<uint4 *>(Xa + ((size_t)(i >> 8) * 2 + s) * pk::SEG +
                                                 (size_t)(i & 255) * pk::STAGE);
            p[0] = make_uint4((uint32_t)v[0], (uint32_t)(v[0] >> 32), (uint32_t)v[1], (uint32_t)(v[1] >> 32));
            p[1] = make_uint4((uint32_t)v[2], (uint32_t)(v[2] >> 32), (uint32_t)v[3], (uint32_t)(v[3] >> 32));
        }
        return;
    }
    const int q = blockIdx.x & 7;
    const int wa = w0 + PF_TW * ((int)blockIdx.x >> 3);
    const int w = wa + wv;
    const int rb = bs->rb;
    const uint64_t pm = bs->pmb[q];
    int i0 = 0;
    for (int qq = 0; qq < q; ++qq) i0 += popc64(bs->pmb[qq]);
    const int ntb = (rb + 3) >> 2;
    const bool tab = pm != 0 && wa + PF_TW > w1 && wa < wc;  // workgroup-uniform
    if (tab) {
        for (int pr = tid; pr < ntb * PF_TW; pr += 256) {
            const int t = pr / PF_TW, wi = pr - t * PF_TW, ww = wa + wi;
            uint64_t v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int s = 4 * t + k;
                v[k] = s < rb && ww >= w1 && ww < wc ? M[(size_t)bs->slot_row[s] * ld + ww] : 0ull;
            }
            uint64_t cur = 0;
            Tbl[t][0][wi] = 0;
#pragma unroll
            for (int e = 1; e < 16; ++e) {
                cur ^= v[__builtin_ctz(e)];
                Tbl[t][e ^ (e >> 1)][wi] = cur;
            }
        }
        __syncthreads();
    }
    uint64_t v = 0;
    if (((pm >> lane) & 1ull) && w < wc) {
        const int i = i0 + popc64(pm & ((1ull << lane) - 1ull));
        if (w < w1) {
            v = Q[(size_t)(r0 + i) * ldq + w];
        } else {
            uint64_t tw[BLK_WORDS];
#pragma unroll
            for (int k = 0; k < BLK_WORDS; ++k) tw[k] = Tin[(size_t)i * BLK_WORDS + k];
            for (int u = 0; u < ntb; ++u)
                v ^= Tbl[u][(tw[u >> 4] >> (4 * (u & 15))) & 15u][wv];
            Q[(size_t)(r0 + i) * ldq + w] = v;
        }
    }
    const uint64_t o = dev::transpose64(v, lane);
    const size_t col = (size_t)64 * (w - w0) + lane;
    *reinterpret_cast<uint64_t *>(Bt + ((col >> 8) * 2 + (q >> 2)) * pk::SEG + (col & 255) * pk::STAGE +
                                  8 * (q & 3)) = o;
}

// Update tiles of block b that hold the head rows' words of block b+1: row
// tiles 0..2 (active rows 0..767 include the BLK_HEAD head rows) x column
// tiles 2 and 3 (words w0+8 .. w0+15).  They run first, in a launch of their
// own, so that the selection of block b+1 can start with the next launch.
__device__ __forceinline__ int head_rows_tiles(int RT) { return min(3, RT); }
__device__ __forceinline__ int head_col_tiles(int CT, bool select_next) {
    return select_next && CT > 2 ? min(4, CT) - 2 : 0;
}
constexpr int HEAD_TILES_MAX = 6;

// Trailing update of block b on the matrix cores: for every row of the update
// (active rows in list order, then for RREF the pivot rows found before the
// block), words w0 .. wc-1 ^= (row's block words) x P~, K = 512.
// head = 1: workgroup i computes head-slice tile i only.  head = 0: workgroup
// 0 selects block b+1 (if any) while workgroups 1.. walk every other tile
// (persistent, row tile fastest).  16 waves per workgroup, each 64 x 64 of
// the 256 x 256 tile (2 x 2 MFMA tiles; four waves per SIMD, so one wave's
// epilogue runs beside another's MFMAs).  Accumulators start at 2^23: every
// count (< 2^10) then sits in the low mantissa bits, so bit 0 of the float is
// its parity; each lane packs its 64 parities (register r of lane l: row
// 8(r/4) + 4(l/32) + r%4, column l%32 of its tile) into one word, one
// v_alignbit per bit, and the in-wave 64 x 64 bit transpose gives lane j the
// 64 lanes' bit j, i.e. 32 columns of one row in each half; lanes j and j^16
// (the two column tiles) combine their halves into the row's word.  Thread t
// then updates word t%4 of tile row t/4: active rows' words inside the block
// are only checked (they must vanish), past it written; pivot rows' every
// word.  The word a thread updates is loaded before the tile's MFMAs, its
// row's active-list entry one tile ahead.
__global__ void __launch_bounds__(BS_THREADS)
k_block_step(const uint8_t *__restrict__ Xa, const uint8_t *__restrict__ Bt, uint64_t *__restrict__ M,
             size_t ld, const uint32_t *__restrict__ act, uint64_t *__restrict__ Q, size_t ldq,
             ElimMeta *__restrict__ meta, int *__restrict__ pivcols, BlockShared *__restrict__ sh,
             BlockState *__restrict__ bst, uint64_t *__restrict__ Tout, int b, int nblk, int wc,
             int rref, int rmax, int head) {
    __shared__ BlockLDS L;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const BlockState *bs = bst + (b & 1);
    const bool select_next = b + 1 < nblk;
    const bool active = bs->active != 0;
    const int w0 = b * BLK_WORDS, w1 = min(w0 + BLK_WORDS, wc);
    const int n_new = bs->n_new;
    const int rows = active ? n_new + (rref ? bs->r0 : 0) : 0;
    const int RT = (rows + 255) >> 8, CT = (wc - w0 + 3) >> 2;
    const int hrt = head_rows_tiles(RT), hct = head_col_tiles(CT, select_next);
    if (!head && blockIdx.x == 0) {
        if (select_next)
            block_select(L, M, ld, act, meta, pivcols, sh, bst + ((b + 1) & 1), Tout, Q, ldq, b + 1,
                         wc, rmax);
        return;
    }
    if (!active) return;
    uint64_t(*Ct)[4] = reinterpret_cast<uint64_t(*)[4]>(&L.H[0][0]);
    const int ntiles = head ? min((int)blockIdx.x + 1, hrt * hct) : RT * CT;
    const int G = head ? 1 : (int)gridDim.x - 1;
    const int t_first = head ? (int)blockIdx.x : (int)blockIdx.x - 1;
    const int wr = wave >> 2, wcl = wave & 3;
    const int half = lane >> 5, l32 = lane & 31;
    constexpr int SEG16 = pk::SEG / 16;
    const int trow = tid >> 2, tk = tid & 3;  // the word this thread updates
    // tile t: the t-th head-slice tile (head launch), else row tile t % RT of
    // column tile t / RT
    auto tile_rc = [&](int t, int &rt, int &ct) {
        if (head) {
            rt = t % hrt;
            ct = 2 + t / hrt;
        } else {
            rt = t % RT;
            ct = t / RT;
        }
    };
    auto is_head = [&](int rt, int ct) { return !head && rt < hrt && ct >= 2 && ct < 2 + hct; };
    auto act_of = [&](int t) -> uint32_t {
        if (t >= ntiles) return 0u;
        int rt, ct;
        tile_rc(t, rt, ct);
        const int row = rt * 256 + trow;
        return row < n_new ? act[row] : 0u;
    };
    bool redo = false, alive = false;
    uint32_t arow = act_of(t_first);
    for (int t = t_first; t < ntiles; t += G) {
        int rt, ct;
        tile_rc(t, rt, ct);
        if (is_head(rt, ct)) {
            arow = act_of(t + G);
            continue;
        }
        const int row = rt * 256 + trow;
        const int w = w0 + 4 * ct + tk;
        uint64_t *dst = nullptr;
        uint64_t old = 0;
        if (row < rows && w < wc) {
            dst = (row < n_new ? M + (size_t)arow * ld : Q + (size_t)(row - n_new) * ldq) + w;
            old = *dst;
        }
        const uint4 *pa = reinterpret_cast<const uint4 *>(Xa + (size_t)rt * 2 * pk::SEG) +
                          (wr * 64 + l32) * 2 + half;
        const uint4 *pb = reinterpret_cast<const uint4 *>(Bt + (size_t)ct * 2 * pk::SEG) +
                          (wcl * 64 + l32) * 2 + half;
        arow = act_of(t + G);
        pk::v16f acc[2][2];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[x][y][r] = 8388608.0f;
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            uint4 fa[2], fb[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                fa[k] = pa[st * SEG16 + k * 64];
                fb[k] = pb[st * SEG16 + k * 64];
            }
            pk::mm<0>(fa, fb, acc);
            pk::mm<1>(fa, fb, acc);
            pk::mm<2>(fa, fb, acc);
            pk::mm<3>(fa, fb, acc);
        }
        {
            uint32_t xh[2] = {0u, 0u};
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        xh[q] = __builtin_amdgcn_alignbit(__float_as_uint(acc[q][h][r]), xh[q], 1);
            const uint64_t y = dev::transpose64(((uint64_t)xh[1] << 32) | xh[0], lane);
            const uint64_t yp = dev::shfl_xor64(y, 16);
            // lane j: tile row j/32, column tile (j/16)%2, r = j%16
            const int tri = lane >> 5, tci = (lane >> 4) & 1, r = lane & 15;
            const uint64_t word = tci == 0 ? ((y & 0xffffffffull) | (yp << 32))
                                           : ((yp >> 32) | (y & 0xffffffff00000000ull));
            Ct[wr * 64 + tri * 32 + 8 * (r >> 2) + (r & 3) + 4 * tci][wcl] = word;
        }
        lds_barrier();
        if (dst) {
            const uint64_t bits = Ct[trow][tk];
            const uint64_t nv = old ^ bits;
            if (row < n_new) {
                if (w < w1) {
                    redo |= nv != 0ull;
                } else {
                    alive |= nv != 0ull;
                    if (bits) *dst = nv;
                }
            } else if (bits) {
                *dst = nv;
            }
        }
        lds_barrier();
    }
    if (__syncthreads_or(redo) && tid == 0) sh->redo = 1;
    if (__syncthreads_or(alive) && tid == 0) atomicMax(&sh->alive, b + 1);
}

}  // namespace

size_t block_scratch_bytes(size_t m, size_t n, bool rref) {
    const size_t wc = words64(n), rmax = m < n ? m : n;
    const size_t rows = m + (rref ? rmax : 0);
    const size_t xa = (rows + 255) / 256 * 2 * pk::SEG;
    const size_t bt = (wc + 3) / 4 * 2 * pk::SEG;
    return BLOCK_T_OFF + (size_t)BLK_COLS * BLK_WORDS * 8 + xa + bt;
}

mela_status enqueue_blocks(mela_ctx *ctx, uint64_t *Mw, size_t ld, uint32_t *act, ElimMeta *meta,
                           int *pivcols, uint64_t *Q, void *blk, size_t m, size_t n, bool rref,
                           size_t b_begin, size_t b_end, bool start) {
    hipStream_t s = ctx->stream;
    const size_t wc = words64(n), rmax = m < n ? m : n;
    const size_t nblk = (wc + BLK_WORDS - 1) / BLK_WORDS;
    char *base = (char *)blk;
    BlockShared *sh = (BlockShared *)base;
    BlockState *bst = (BlockState *)(base + BLOCK_STATE_OFF);
    uint64_t *T = (uint64_t *)(base + BLOCK_T_OFF);
    const size_t rows = m + (rref ? rmax : 0);
    uint8_t *Xa = (uint8_t *)(base + BLOCK_T_OFF + (size_t)BLK_COLS * BLK_WORDS * 8);
    uint8_t *Bt = Xa + (rows + 255) / 256 * 2 * pk::SEG;
    const int cu = ctx->cu_count > 0 ? ctx->cu_count : 256;
    const int xa_blocks = (int)std::min<size_t>((2 * ((rows + 255) & ~(size_t)255) + 255) / 256, 1024);
    auto pform = [&](size_t b) {
        const size_t w0 = b * BLK_WORDS;
        const int ntw = (int)((wc - w0 + PF_TW - 1) / PF_TW);
        hipLaunchKernelGGL(k_block_pform, dim3((unsigned)(8 * ntw + xa_blocks)), dim3(256), 0, s,
                           (const uint64_t *)Mw, ld, act, (const BlockShared *)sh,
                           (const BlockState *)(bst + (b & 1)), (const uint64_t *)T, Q, ld, Xa, Bt,
                           (int)b, (int)wc, ntw, rref ? 1 : 0, xa_blocks);
    };
    if (start && b_begin == 0) {
        hipLaunchKernelGGL(k_block_init, dim3(1), dim3(64), 0, s, sh, bst);
        hipLaunchKernelGGL(k_block_select, dim3(1), dim3(BS_THREADS), 0, s, (const uint64_t *)Mw, ld,
                           (const uint32_t *)act, meta, pivcols, sh, bst, T, Q, ld, (int)wc, (int)rmax);
        pform(0);
    }
    for (size_t b = b_begin; b < b_end; ++b) {
        // the head slice of block b's update, then the rest of it beside the
        // selection of block b+1
        if (b + 1 < nblk)
            hipLaunchKernelGGL(k_block_step, dim3(HEAD_TILES_MAX), dim3(BS_THREADS), 0, s,
                               (const uint8_t *)Xa, (const uint8_t *)Bt, Mw, ld, (const uint32_t *)act, Q,
                               ld, meta, pivcols, sh, bst, T, (int)b, (int)nblk, (int)wc, rref ? 1 : 0,
                               (int)rmax, 1);
        hipLaunchKernelGGL(k_block_step, dim3((unsigned)cu), dim3(BS_THREADS), 0, s, (const uint8_t *)Xa,
                           (const uint8_t *)Bt, Mw, ld, (const uint32_t *)act, Q, ld, meta, pivcols, sh, bst,
                           T, (int)b, (int)nblk, (int)wc, rref ? 1 : 0, (int)rmax, 0);
        if (b + 1 < nblk) pform(b + 1);
    }
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

}  // namespace mela
