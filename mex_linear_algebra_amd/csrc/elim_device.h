// Workgroup-level building blocks of GF(2) elimination, shared by the
// single-matrix kernels (elim.hip) and the batched persistent kernel
// (batch.hip).
#pragma once

#include "common.h"

namespace mela {

// Device state of one single-matrix elimination (elim.hip, elim_block.hip);
// the pivot columns follow it in the same buffer.
struct ElimMeta {
    int n_act;  // active (not yet pivot) rows
    int r;      // pivot rows found so far
    int r0;     // pivot rows before the current panel
    int rho;    // pivots in the current panel
    uint64_t pm;        // pivot bits (columns 64g + c) of the current panel
    int finished;       // every active row is zero: the remaining panels have no pivot
    int last_swept;     // last panel with pivots (its sweep ran), -1 before the first
    uint64_t pm_prev;   // pm and r0 of the previous panel (RREF: its word update is
    int r0_prev;        // applied by the next panel's sweep)
    int done_at;        // first panel whose selection found the elimination complete
    int sel_row[64];           // selected row ids (selection slot -> row)
    uint64_t comb[64];  // by pivot bit c: selection slots XORed into pivot row c
};

namespace dev {

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }
__device__ __forceinline__ int ffs64(uint64_t x) { return __ffsll((unsigned long long)x) - 1; }

// XOR of v over the 64 lanes of the wave (every lane receives the result;
// all 64 lanes must be active).  DPP row shifts give the inclusive XOR of
// each 16-lane row, two row broadcasts carry rows 0-2 into lane 63, whose
// value is read out: six VALU steps per half instead of six LDS permutes.
__device__ __forceinline__ uint32_t wave_xor32(uint32_t v) {
    v ^= __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xF, 0xF, true);   // row_shr:1
    v ^= __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xF, 0xF, true);   // row_shr:2
    v ^= __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, true);   // row_shr:4
    v ^= __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xF, 0xF, true);   // row_shr:8
    v ^= __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 -> rows 1, 3
    v ^= __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 -> rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ uint64_t wave_xor64(uint64_t v) {
    const uint32_t lo = wave_xor32((uint32_t)v), hi = wave_xor32((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// OR of v over the 64 lanes of the wave, the same DPP pattern as wave_xor32
// (all 64 lanes must be active).
__device__ __forceinline__ uint32_t wave_or32(uint32_t v) {
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xF, 0xF, true);   // row_shr:1
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xF, 0xF, true);   // row_shr:2
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, true);   // row_shr:4
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xF, 0xF, true);   // row_shr:8
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 -> rows 1, 3
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 -> rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
}

// Wave-level Gauss-Jordan over the 64-bit words v (one per lane; lanes with
// valid == false take no part).  Afterwards pivot lanes hold the RREF basis of
// span{v} (pbit = pivot bit = lowest set bit of the lane's word), non-pivot
// lanes hold 0; comb (if tracked) is the set of original lanes XORed into each
// lane.  Pivot of column c = the lowest non-pivot lane with bit c set.  The
// pivot set is kept as a wave-uniform mask so each column costs one ballot,
// one s_ff1, two (four) v_readlane and an exec-masked XOR.
// A row XOR never sets a bit in a column that no row has, so only the
// columns of the rows' OR can take pivots; with fewer than this many of them
// (a sparse panel word) the elimination visits those columns alone.
constexpr int GJ_SPARSE_COLS = 48;

template <bool TRACK>
__device__ __forceinline__ void wave_gj(uint64_t &v, uint64_t &comb, bool valid, bool &piv,
                                        int &pbit, int lane) {
    uint32_t lo = valid ? (uint32_t)v : 0u, hi = valid ? (uint32_t)(v >> 32) : 0u;
    uint32_t clo = (uint32_t)comb, chi = (uint32_t)(comb >> 32);
    uint64_t pivmask = 0;  // wave-uniform
    auto column = [&](int c) {
        const uint32_t word = c < 32 ? lo : hi;
        const bool has = (word >> (c & 31)) & 1u;
        const uint64_t hb = __ballot(has);
        const uint64_t cand = hb & ~pivmask;
        if (cand == 0) return;
        const int f = ffs64(cand);
        const uint32_t plo = __builtin_amdgcn_readlane(lo, f);
        const uint32_t phi = __builtin_amdgcn_readlane(hi, f);
        uint32_t pclo = 0, pchi = 0;
        if (TRACK) {
            pclo = __builtin_amdgcn_readlane(clo, f);
            pchi = __builtin_amdgcn_readlane(chi, f);
        }
        pivmask |= 1ull << f;
        if (has && lane != f) {
            lo ^= plo;
            hi ^= phi;
            if (TRACK) {
                clo ^= pclo;
                chi ^= pchi;
            }
        }
    };
    const uint64_t cols = ((uint64_t)wave_or32(hi) << 32) | wave_or32(lo);
    if (popc64(cols) >= GJ_SPARSE_COLS) {
#pragma unroll
        for (int c = 0; c < 64; ++c) column(c);
    } else {
        for (uint64_t cs = cols; cs; cs &= cs - 1) column(ffs64(cs));
    }
    v = ((uint64_t)hi << 32) | lo;
    if (TRACK) comb = ((uint64_t)chi << 32) | clo;
    piv = (pivmask >> lane) & 1ull;
    pbit = piv ? ffs64(v) : -1;
}

// Bit compress (x's bits at the set positions of mask m, packed to the low
// end in order) by the six-step parallel-suffix method of Hacker's Delight
// (2nd ed., section 7-4, "compress").  The six move masks depend on m only, so
// a null-basis extraction computes them once per column word and applies them
// to every row word of that column.
__host__ __device__ __forceinline__ void compress_masks(uint64_t m, uint64_t mv[6]) {
    uint64_t mk = ~m << 1;  // bit j: count of zeros of m below j, parity by prefix XOR
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        uint64_t mp = mk ^ (mk << 1);
        mp ^= mp << 2;
        mp ^= mp << 4;
        mp ^= mp << 8;
        mp ^= mp << 16;
        mp ^= mp << 32;
        const uint64_t v = mp & m;  // bits that move right by 2^i in this step
        mv[i] = v;
        m = (m ^ v) | (v >> (1 << i));
        mk &= ~mp;
    }
}

__host__ __device__ __forceinline__ uint64_t compress_apply(uint64_t x, uint64_t m, const uint64_t mv[6]) {
    x &= m;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const uint64_t t = x & mv[i];
        x = (x ^ t) | (t >> (1 << i));
    }
    return x;
}

// Pivots the first 64 active rows must give for the selection to complete
// them by single insertions instead of the general chunked path.
constexpr int FAST_MIN_PIVOTS = 48;

// LDS state of one panel selection.
template <int NT>
struct SelectLDS {
    uint64_t bas[64];  // RREF basis of the panel, indexed by pivot bit
    uint64_t cy[NT];   // candidate words of the current chunk
    int cpos[NT];      // their positions in the active list
    int wcnt[NT / 64];
    int selpos[64];
    int selrow[64];
    int flag[64];
    uint32_t mv[64];
    unsigned long long pm;
    int nsel;
    int fast;   // the fast path selected 64 rows (combinations already final)
    int added;  // the last candidate group added pivots
};

// Result of a panel selection (valid on every thread after the call).
struct PanelOut {
    uint64_t pm;        // pivot bits of the panel
    int rho;            // number of pivots (= popcount(pm))
    int n_act_new;      // active rows left
};

// Selects <= 64 active rows whose panel words (word g of each row) span all
// active rows' panel words, removes them from the active list act[0..n_act)
// (stable for the rest of the list up to hole filling), and returns for each
// pivot bit c the RREF basis combination comb[c] over the selected rows
// selrow[0..rho).  All NT threads of the workgroup must call it.
//   row(q)   : pointer to the row of active-list entry q's row id
//   act      : active list (global or LDS), ACT = its element type
//   pre      : if have_pre, thread t's panel word of active-list entry t
//              (t < min(n_act, NT)), already loaded by the caller
//   prew     : optional panel words of all active-list entries, in list
//              order (read instead of each row's word g past the first NT)
//   actc     : optional LDS copy of act (ACT_HEAD + 64 entries) for a global
//              act: waves 1-4 fill it with act[0 .. ACT_HEAD) and the last 64
//              entries while wave 0 runs the fast path, so the final
//              reordering reads the selected and the moved entries from LDS
constexpr int ACT_HEAD = 192;  // positions the fast path can select (64 + 128)
template <int NT, typename ACT, typename RowFn>
__device__ PanelOut wg_panel_select(SelectLDS<NT> &L, ACT *act, int n_act, int g, RowFn row,
                                    uint64_t *comb_out, int *selrow_out, bool have_pre = false,
                                    uint64_t pre = 0, uint32_t *actc = nullptr,
                                    const uint64_t *prew = nullptr) {
    // the copy needs waves 1-4 (workgroups of at least 320 threads)
    if (NT < 64 + ACT_HEAD + 64) actc = nullptr;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < 64) L.bas[tid] = 0;
    if (tid == 0) {
        L.pm = 0;
        L.nsel = 0;
        L.fast = 0;
    }
    __syncthreads();
    // ---- fast path (dense panels): one tracked Gauss-Jordan over the first 64
    // active rows, then single insertions of the next rows until 64 pivots; a
    // panel that does not reach 64 pivots this way takes the general path.
    const int n0 = n_act < NT ? n_act : NT;
    if (n_act >= 64) {
        if (tid < n0) L.cy[tid] = have_pre ? pre : row((int)act[tid])[g];
        __syncthreads();
        // the rank of the first 64 words is at most the number of nonzero
        // ones: with fewer than FAST_MIN_PIVOTS the general path runs at once
        if (wave == 0 && popc64(__ballot(L.cy[lane] != 0ull)) >= FAST_MIN_PIVOTS) {
            uint64_t v = L.cy[lane];
            uint64_t comb = 1ull << lane;
            bool piv;
            int pbit;
            wave_gj<true>(v, comb, true, piv, pbit, lane);
            uint64_t pivl = __ballot(piv);  // pivot lanes = selection slots
            int slotpos = lane;              // active-list position of slot `lane`
            // single insertions only when the first 64 rows came close to 64
            // pivots (sparse panels would scan 128 rows and then fail)
            const int lim = popc64(pivl) < FAST_MIN_PIVOTS ? 0 : n0 < ACT_HEAD ? n0 : ACT_HEAD;
            for (int pos = 64; pos < lim && pivl != ~0ull; ++pos) {
                const uint64_t x = L.cy[pos];
                if (x == 0) continue;
                const bool use = piv && ((x >> pbit) & 1ull);
                const uint64_t red = wave_xor64(use ? v : 0ull);
                const uint64_t cred = wave_xor64(use ? comb : 0ull);
                const uint64_t z = x ^ red;  // x reduced by the RREF basis
                if (z == 0) continue;
                const int sl = ffs64(~pivl);  // a free slot
                const uint64_t cz = cred ^ (1ull << sl);
                const int p = ffs64(z);
                if (piv && ((v >> p) & 1ull)) {
                    v ^= z;
                    comb ^= cz;
                }
                if (lane == sl) {
                    v = z;
                    comb = cz;
                    piv = true;
                    pbit = p;
                    slotpos = pos;
                }
                pivl |= 1ull << sl;
            }
            if (pivl == ~0ull) {  // 64 pivots: slot = lane
                L.selpos[lane] = slotpos;
                comb_out[pbit] = comb;
                if (lane == 0) {
                    L.pm = ~0ull;
                    L.nsel = 64;
                    L.fast = 1;
                }
            }
        } else if (actc && tid >= 64 && tid < 64 + ACT_HEAD + 64) {
            const int k = tid - 64;
            const int p = k < ACT_HEAD ? k : n_act - 64 + (k - ACT_HEAD);
            if (p < n_act) actc[k] = (uint32_t)act[p];
        }
        __syncthreads();
    }
    // the LDS copy of act is filled when the block above ran (n_act >= 64)
    const bool have_actc = actc != nullptr && n_act >= 64;
    // chunks of NT active rows; the next chunk's words are loaded while the
    // current chunk is reduced (positions past the end re-read a valid row)
    uint64_t xnext = 0;
    if (!L.fast && n_act > 0)
        xnext = have_pre && tid < n_act ? pre : row((int)act[tid < n_act ? tid : n_act - 1])[g];
    for (int base = 0; base < n_act && !L.fast; base += NT) {
        const int pos = base + tid;
        uint64_t x = pos < n_act ? xnext : 0ull;
        {
            const int pn = pos + NT;
            const int pc = pn < n_act ? pn : n_act - 1;
            xnext = prew ? prew[pc] : row((int)act[pc])[g];
        }
        const uint64_t pm = L.pm;
        uint64_t y = x;
        uint64_t hit = x & pm;
        while (hit) {
            const int c = ffs64(hit);
            hit &= hit - 1;
            y ^= L.bas[c];
        }
        const uint64_t bal = __ballot(y != 0);
        if (lane == 0) L.wcnt[wave] = popc64(bal);
        __syncthreads();
        int off = 0, ncand = 0;
        for (int q = 0; q < NT / 64; ++q) {
            const int cq = L.wcnt[q];
            if (q < wave) off += cq;
            ncand += cq;
        }
        if (y != 0) {
            const int my = off + popc64(bal & ((1ull << lane) - 1ull));
            L.cy[my] = y;
            L.cpos[my] = pos;
        }
        __syncthreads();
        // groups of 64 candidates through wave 0's Gauss-Jordan; after a group
        // that added pivots every thread reduces the remaining candidates by
        // the basis as it is now and compacts them (once the panel's pivots
        // are found the rest of the chunk usually vanishes here, instead of
        // being reduced 64 at a time by wave 0)
        for (int gb = 0;;) {
            // every thread reads the loop state before the barrier; wave 0
            // writes it only after
            const bool more = gb < ncand && L.pm != ~0ull;
            __syncthreads();
            if (!more) break;
            if (wave == 0) {
                uint64_t pmw = L.pm;
                int nsel = L.nsel;
                const int ci = gb + lane;
                const bool valid = ci < ncand;
                uint64_t v = valid ? L.cy[ci] : 0;
                uint64_t h2 = v & pmw;  // pivots added since the last filter
                while (h2) {
                    const int c = ffs64(h2);
                    h2 &= h2 - 1;
                    v ^= L.bas[c];
                }
                uint64_t dummy = 0;
                bool piv;
                int pbit;
                wave_gj<false>(v, dummy, valid && v != 0, piv, pbit, lane);
                const uint64_t pl = __ballot(piv);
                uint64_t newmask = 0;
                if (piv) L.bas[pbit] = v;
                __threadfence_block();
                for (uint64_t t = pl; t; t &= t - 1)
                    newmask |= 1ull << __builtin_amdgcn_readlane(pbit, ffs64(t));
                // the new vectors are RREF among themselves and zero on the old
                // pivot bits: one pass clears the new bits from the old basis
                if ((pmw >> lane) & 1ull) {
                    uint64_t b = L.bas[lane];
                    uint64_t hb = b & newmask;
                    while (hb) {
                        const int c = ffs64(hb);
                        hb &= hb - 1;
                        b ^= L.bas[c];
                    }
                    L.bas[lane] = b;
                }
                if (piv) L.selpos[nsel + popc64(pl & ((1ull << lane) - 1ull))] = L.cpos[ci];
                __threadfence_block();
                nsel += popc64(pl);
                pmw |= newmask;
                if (lane == 0) {
                    L.pm = pmw;
                    L.nsel = nsel;
                    L.added = pl != 0;
                }
            }
            __syncthreads();
            gb += 64;
            if (L.added && gb < ncand && L.pm != ~0ull) {
                const int ci = gb + tid;
                uint64_t yy = 0;
                int cp = 0;
                if (ci < ncand) {
                    yy = L.cy[ci];
                    cp = L.cpos[ci];
                    const uint64_t pmn = L.pm;
                    for (uint64_t h = yy & pmn; h; h &= h - 1) yy ^= L.bas[ffs64(h)];
                }
                const uint64_t b2 = __ballot(yy != 0);
                if (lane == 0) L.wcnt[wave] = popc64(b2);
                __syncthreads();
                int off2 = 0, n2 = 0;
                for (int q = 0; q < NT / 64; ++q) {
                    const int cq = L.wcnt[q];
                    if (q < wave) off2 += cq;
                    n2 += cq;
                }
                if (yy != 0) {
                    const int my = gb + off2 + popc64(b2 & ((1ull << lane) - 1ull));
                    L.cy[my] = yy;
                    L.cpos[my] = cp;
                }
                ncand = gb + n2;
                __syncthreads();
            }
        }
        if (L.pm == ~0ull) break;
    }

    PanelOut out;
    const int nsel = L.nsel;
    const int n_new = n_act - nsel;
    if (wave == 0) {
        // holes (selected positions below n_new) take the unselected rows of
        // positions n_new .. n_act-1
        int my_pos = -1;
        if (lane < nsel) {
            my_pos = L.selpos[lane];
            L.selrow[lane] = have_actc && my_pos < ACT_HEAD ? (int)actc[my_pos] : (int)act[my_pos];
        }
        L.flag[lane] = 0;
        __threadfence_block();
        if (lane < nsel && my_pos >= n_new) L.flag[my_pos - n_new] = 1;
        __threadfence_block();
        const bool is_hole = lane < nsel && my_pos < n_new;
        const bool is_mover = lane < nsel && L.flag[lane] == 0;
        const uint64_t hb = __ballot(is_hole), mb = __ballot(is_mover);
        const uint64_t below = (1ull << lane) - 1ull;
        // position n_new + lane is tail entry lane + 64 - nsel of the copy
        if (is_mover)
            L.mv[popc64(mb & below)] =
                have_actc ? actc[ACT_HEAD + lane + 64 - nsel] : (uint32_t)act[n_new + lane];
        __threadfence_block();
        if (is_hole) act[my_pos] = (ACT)L.mv[popc64(hb & below)];
        if (lane < nsel) selrow_out[lane] = L.selrow[lane];
        if (!L.fast) {
            // RREF of the selected rows' panel words, tracking combinations
            uint64_t v = 0, comb = 0;
            if (lane < nsel) {
                v = prew ? prew[L.selpos[lane]] : row(L.selrow[lane])[g];
                comb = 1ull << lane;
            }
            bool piv;
            int pbit;
            wave_gj<true>(v, comb, lane < nsel, piv, pbit, lane);
            if (piv) comb_out[pbit] = comb;
        }
    }
    __syncthreads();
    out.pm = L.pm;
    out.rho = nsel;
    out.n_act_new = n_new;
    return out;
}

}  // namespace dev
}  // namespace mela
