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
// time, with three launches:
//   k_panel_select  one workgroup finds <= 64 active rows whose panel words
//                   span all active rows' panel words (wave Gauss-Jordan with
//                   ballot/readlane), removes them from the active list and
//                   computes the RREF basis of the panel plus, for each basis
//                   vector, which selected rows XOR to it;
//   k_build_pivots  forms the new pivot rows Q (full width) from the selected
//                   rows;
//   k_update        clears the panel's pivot columns from every active row
//                   (and, for RREF, from every earlier pivot row) with one
//                   table XOR per 4 pivot columns (M4RI) or one XOR per pivot
//                   row (GE), sweeping words g+1 .. wc-1 -- the HBM-bound part.
#include "common.h"

namespace mela {

struct ElimMeta {
    int n_act;  // active (not yet pivot) rows
    int r;      // pivot rows found so far
    int r0;     // pivot rows before the current panel
    int rho;    // pivots in the current panel
    unsigned long long pm;     // pivot bits (columns 64g + c) of the current panel
    int sel_row[64];           // selected row ids (selection slot -> row)
    unsigned long long comb[64];  // by pivot bit c: selection slots XORed into pivot row c
};

namespace {

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// Wave-level Gauss-Jordan over the 64-bit words v (one per lane, lanes with
// valid == false take no part).  Afterwards pivot lanes hold the RREF basis of
// span{v} (pbit = pivot bit), non-pivot lanes hold 0; comb (if tracked) is the
// set of original lanes XORed into each lane.  Pivot of column c = the lowest
// non-pivot lane with bit c set.
template <bool TRACK>
__device__ __forceinline__ void wave_gj(uint64_t &v, uint64_t &comb, bool valid, bool &piv,
                                        int &pbit, int lane) {
    piv = false;
    pbit = -1;
    if (!valid) v = 0;
    for (int c = 0; c < 64; ++c) {
        const bool has = ((v >> c) & 1ull) != 0;
        const uint64_t bal = __ballot(has && !piv);
        if (bal == 0) continue;
        const int f = __ffsll((unsigned long long)bal) - 1;
        const uint64_t pv = readlane64(v, f);
        uint64_t pc = 0;
        if (TRACK) pc = readlane64(comb, f);
        if (lane == f) {
            piv = true;
            pbit = c;
        } else if (has) {
            v ^= pv;
            if (TRACK) comb ^= pc;
        }
    }
}

constexpr int PS_THREADS = 1024;

// One workgroup.  Streams the active rows' panel words in chunks of 1024,
// filters each against the current RREF basis, and inserts the survivors 64 at
// a time with a wave Gauss-Jordan until 64 pivots are found or rows run out.
__global__ void __launch_bounds__(PS_THREADS)
k_panel_select(const uint64_t *__restrict__ M, size_t ld, uint32_t *__restrict__ act,
               ElimMeta *__restrict__ meta, int *__restrict__ pivcols, int g, int rmax) {
    __shared__ uint64_t s_bas[64];   // RREF basis by pivot bit
    __shared__ uint64_t s_cy[PS_THREADS];
    __shared__ int s_cpos[PS_THREADS];
    __shared__ int s_wcnt[16];
    __shared__ int s_selpos[64];
    __shared__ int s_selrow[64];
    __shared__ int s_flag[64];
    __shared__ unsigned long long s_pm;
    __shared__ int s_nsel;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n_act = meta->n_act;
    const int r_before = meta->r;
    if (tid < 64) s_bas[tid] = 0;
    if (tid == 0) {
        s_pm = 0;
        s_nsel = 0;
    }
    __syncthreads();

    const int room = rmax - r_before;  // pivots still possible
    for (int base = 0; base < n_act && room > 0; base += PS_THREADS) {
        const int pos = base + tid;
        uint64_t x = 0;
        if (pos < n_act) x = M[(size_t)act[pos] * ld + g];
        // filter against the current basis (one pass is exact for RREF)
        uint64_t pm = s_pm;
        uint64_t y = x;
        uint64_t hit = x & pm;
        while (hit) {
            const int c = __ffsll((unsigned long long)hit) - 1;
            hit &= hit - 1;
            y ^= s_bas[c];
        }
        const uint64_t bal = __ballot(y != 0);
        if (lane == 0) s_wcnt[wave] = popc64(bal);
        __syncthreads();
        int off = 0, ncand = 0;
        for (int q = 0; q < 16; ++q) {
            const int cq = s_wcnt[q];
            if (q < wave) off += cq;
            ncand += cq;
        }
        if (y != 0) {
            const int my = off + popc64(bal & ((1ull << lane) - 1ull));
            s_cy[my] = y;
            s_cpos[my] = pos;
        }
        __syncthreads();
        if (wave == 0) {
            uint64_t pmw = s_pm;
            int nsel = s_nsel;
            for (int gb = 0; gb < ncand && pmw != ~0ull; gb += 64) {
                const int ci = gb + lane;
                const bool valid = ci < ncand;
                uint64_t v = valid ? s_cy[ci] : 0;
                // reduce against pivots added since the chunk filter
                uint64_t h2 = v & pmw;
                while (h2) {
                    const int c = __ffsll((unsigned long long)h2) - 1;
                    h2 &= h2 - 1;
                    v ^= s_bas[c];
                }
                uint64_t dummy = 0;
                bool piv;
                int pbit;
                wave_gj<false>(v, dummy, valid && v != 0, piv, pbit, lane);
                // (pivots can never exceed min(m,n): every pivot consumes an
                // active row and a column below n)
                const uint64_t pl = __ballot(piv);
                // clear new pivot bits from the old basis (new vectors are RREF
                // among themselves and zero on old pivot bits)
                uint64_t newmask = 0;
                // gather (pbit, vector) of new pivots through LDS
                if (piv) s_bas[pbit] = v;
                __threadfence_block();
                for (uint64_t t = pl; t; t &= t - 1) {
                    const int l = __ffsll((unsigned long long)t) - 1;
                    newmask |= 1ull << __builtin_amdgcn_readlane(pbit, l);
                }
                if (lane < 64 && ((pmw >> lane) & 1ull)) {
                    uint64_t b = s_bas[lane];
                    uint64_t hb = b & newmask;
                    while (hb) {
                        const int c = __ffsll((unsigned long long)hb) - 1;
                        hb &= hb - 1;
                        b ^= s_bas[c];
                    }
                    s_bas[lane] = b;
                }
                // record selected positions in candidate order
                if (piv) {
                    const int slot = nsel + popc64(pl & ((1ull << lane) - 1ull));
                    s_selpos[slot] = s_cpos[ci];
                }
                __threadfence_block();
                nsel += popc64(pl);
                pmw |= newmask;
            }
            if (lane == 0) {
                s_pm = pmw;
                s_nsel = nsel;
            }
        }
        __syncthreads();
        if (s_pm == ~0ull) break;
    }

    // ---- finalise on wave 0: selected rows, act compaction, RREF + comb ----
    if (wave != 0) return;
    const int nsel = s_nsel;
    const int n_new = n_act - nsel;
    int my_pos = -1;
    if (lane < nsel) {
        my_pos = s_selpos[lane];
        s_selrow[lane] = (int)act[my_pos];
    }
    s_flag[lane] = 0;
    __threadfence_block();
    if (lane < nsel && my_pos >= n_new) s_flag[my_pos - n_new] = 1;
    __threadfence_block();
    // holes: selected positions below n_new; movers: unselected positions >= n_new
    const bool is_hole = lane < nsel && my_pos < n_new;
    const bool is_mover = lane < nsel && s_flag[lane] == 0;
    const uint64_t hb = __ballot(is_hole), mb = __ballot(is_mover);
    const uint64_t below = (1ull << lane) - 1ull;
    int mover_rank = is_mover ? popc64(mb & below) : -1;
    uint32_t mover_row = is_mover ? act[n_new + lane] : 0;
    // hole k takes mover k: pass mover rows by rank through LDS
    __shared__ uint32_t s_mv[64];
    if (is_mover) s_mv[mover_rank] = mover_row;
    __threadfence_block();
    if (is_hole) act[my_pos] = s_mv[popc64(hb & below)];

    // RREF basis of the selected rows' original panel words, with combs
    uint64_t v = 0, comb = 0;
    if (lane < nsel) {
        v = M[(size_t)s_selrow[lane] * ld + g];
        comb = 1ull << lane;
    }
    bool piv;
    int pbit;
    wave_gj<true>(v, comb, lane < nsel, piv, pbit, lane);
    if (lane < nsel) meta->sel_row[lane] = s_selrow[lane];
    // pivot bits
    uint64_t pm = 0;
    for (uint64_t t = __ballot(piv); t; t &= t - 1) {
        const int l = __ffsll((unsigned long long)t) - 1;
        pm |= 1ull << __builtin_amdgcn_readlane(pbit, l);
    }
    if (piv) meta->comb[pbit] = comb;
    const int rho = popc64(pm);
    if (lane == 0) {
        meta->r0 = r_before;
        meta->rho = rho;
        meta->pm = pm;
        meta->r = r_before + rho;
        meta->n_act = n_new;
    }
    // pivot columns ascending: bit c is the popcount(pm below c)-th new pivot
    if ((pm >> lane) & 1ull) pivcols[r_before + popc64(pm & below)] = g * 64 + lane;
}

// Q[r0 + j][w] for the j-th new pivot (ascending bit c) = XOR of the selected
// rows in comb[c]; words below g are zero (all active rows are zero there).
__global__ void __launch_bounds__(256)
k_build_pivots(const uint64_t *__restrict__ M, size_t ld, uint64_t *__restrict__ Q, size_t ldq,
               const ElimMeta *__restrict__ meta, int g, int wc) {
    const uint64_t pm = meta->pm;
    const int j = blockIdx.y;  // pivot bit c = the j-th set bit of pm
    if (j >= popc64(pm)) return;
    uint64_t t = pm;
    for (int q = 0; q < j; ++q) t &= t - 1;
    const int c = __ffsll((unsigned long long)t) - 1;
    const uint64_t comb = meta->comb[c];
    const int r0 = meta->r0;
    const int w = blockIdx.x * 256 + threadIdx.x;
    if (w >= wc) return;
    uint64_t acc = 0;
    if (w >= g) {
        for (uint64_t s = comb; s; s &= s - 1) {
            const int slot = __ffsll((unsigned long long)s) - 1;
            acc ^= M[(size_t)meta->sel_row[slot] * ld + w];
        }
    }
    Q[(size_t)(r0 + j) * ldq + w] = acc;
}

// Row pointer of unified row index q: active rows first, then (RREF) the
// earlier pivot rows.
__device__ __forceinline__ uint64_t *row_ptr(uint64_t *M, size_t ld, const uint32_t *act,
                                             uint64_t *Q, size_t ldq, int n_act, int q) {
    return q < n_act ? M + (size_t)act[q] * ld : Q + (size_t)(q - n_act) * ldq;
}

constexpr int UP_TILE = 64;  // words per workgroup (one per lane)

// M4RI form: 16 nibble tables of the panel's pivot rows (slot c = pivot bit c,
// zero where c is not a pivot) in LDS; each row XORs 16 table rows selected
// by the 16 nibbles of its panel word.  Words g+1 .. wc-1 only (word g is
// handled by k_update_panelword so that every tile reads an unmodified x).
__global__ void __launch_bounds__(256)
k_update_m4ri(uint64_t *__restrict__ M, size_t ld, const uint32_t *__restrict__ act,
              uint64_t *__restrict__ Q, size_t ldq, const ElimMeta *__restrict__ meta, int g,
              int wc, int rref) {
    __shared__ uint64_t T[16][16][UP_TILE];
    const uint64_t pm = meta->pm;
    if (pm == 0) return;
    const int n_act = meta->n_act, r0 = meta->r0;
    const int total = n_act + (rref ? r0 : 0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int w = g + 1 + blockIdx.x * UP_TILE + lane;
    const bool wok = w < wc;
    if (blockIdx.y * 4 >= (unsigned)total) return;
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
        const int t = wave * 4 + tt;
        uint64_t q[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int c = 4 * t + b;
            q[b] = 0;
            if (wok && ((pm >> c) & 1ull))
                q[b] = Q[(size_t)(r0 + popc64(pm & ((1ull << c) - 1ull))) * ldq + w];
        }
        uint64_t L[16];
        L[0] = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int u = 0; u < (1 << b); ++u) L[(1 << b) + u] = L[u] ^ q[b];
#pragma unroll
        for (int v = 0; v < 16; ++v) T[t][v][lane] = L[v];
    }
    __syncthreads();
    const int stride = gridDim.y * 4;
    for (int q0 = blockIdx.y * 4 + wave; q0 < total; q0 += stride) {
        uint64_t *rp = row_ptr(M, ld, act, Q, ldq, n_act, q0);
        const uint64_t x = rp[g] & pm;
        if (x == 0 || !wok) continue;
        uint64_t a = rp[w];
#pragma unroll
        for (int t = 0; t < 16; ++t) a ^= T[t][(x >> (4 * t)) & 15u][lane];
        rp[w] = a;
    }
}

// GE form: the panel's pivot rows themselves in LDS; each row XORs one pivot
// row per set pivot bit of its panel word.
__global__ void __launch_bounds__(256)
k_update_ge(uint64_t *__restrict__ M, size_t ld, const uint32_t *__restrict__ act,
            uint64_t *__restrict__ Q, size_t ldq, const ElimMeta *__restrict__ meta, int g, int wc,
            int rref) {
    __shared__ uint64_t P[64][UP_TILE];
    const uint64_t pm = meta->pm;
    if (pm == 0) return;
    const int n_act = meta->n_act, r0 = meta->r0;
    const int total = n_act + (rref ? r0 : 0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int w = g + 1 + blockIdx.x * UP_TILE + lane;
    const bool wok = w < wc;
    if (blockIdx.y * 4 >= (unsigned)total) return;
    for (int c = wave; c < 64; c += 4) {
        uint64_t v = 0;
        if (wok && ((pm >> c) & 1ull))
            v = Q[(size_t)(r0 + popc64(pm & ((1ull << c) - 1ull))) * ldq + w];
        P[c][lane] = v;
    }
    __syncthreads();
    const int stride = gridDim.y * 4;
    for (int q0 = blockIdx.y * 4 + wave; q0 < total; q0 += stride) {
        uint64_t *rp = row_ptr(M, ld, act, Q, ldq, n_act, q0);
        uint64_t x = rp[g] & pm;
        if (x == 0 || !wok) continue;
        uint64_t a = rp[w];
        while (x) {
            const int c = __ffsll((unsigned long long)x) - 1;
            x &= x - 1;
            a ^= P[c][lane];
        }
        rp[w] = a;
    }
}

// RREF only: word g of the earlier pivot rows (active rows' word g is dead).
__global__ void __launch_bounds__(256)
k_update_panelword(uint64_t *__restrict__ Q, size_t ldq, const ElimMeta *__restrict__ meta,
                   int g) {
    const uint64_t pm = meta->pm;
    const int r0 = meta->r0;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (pm == 0 || q >= r0) return;
    uint64_t x = Q[(size_t)q * ldq + g];
    uint64_t h = x & pm;
    uint64_t a = x;
    while (h) {
        const int c = __ffsll((unsigned long long)h) - 1;
        h &= h - 1;
        a ^= Q[(size_t)(r0 + popc64(pm & ((1ull << c) - 1ull))) * ldq + g];
    }
    Q[(size_t)q * ldq + g] = a;
}

__global__ void k_elim_init(ElimMeta *meta, uint32_t *act, int m) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < m) act[i] = (uint32_t)i;
    if (i == 0) {
        meta->n_act = m;
        meta->r = 0;
        meta->r0 = 0;
        meta->rho = 0;
        meta->pm = 0;
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

// Z (n x f logical, column-major): row p of column j is 1 if p is the j-th
// free column, bit free[j] of pivot row rowmap[p] if p is a pivot column, 0
// otherwise (gf2null.m:53-55; null_gf2.c:223-273 computes the same entries by
// back-substitution).
__global__ void __launch_bounds__(256)
k_null_logical(const uint64_t *__restrict__ Q, size_t ldq, const int *__restrict__ rowmap,
               const int *__restrict__ freecols, size_t n, size_t f, uint8_t *__restrict__ Z) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int rm = rowmap[p];
    for (size_t j = blockIdx.y; j < f; j += gridDim.y) {
        uint8_t v;
        if (rm >= 0) {
            const int fc = freecols[j];
            v = (uint8_t)((Q[(size_t)rm * ldq + (fc >> 6)] >> (fc & 63)) & 1ull);
        } else {
            v = (size_t)(-rm - 1) == j;
        }
        Z[j * n + p] = v;
    }
}

// Packed Z (n x f, row-major, ldz words): word (p, jw) gathers bits
// freecols[64 jw .. 64 jw + 63] of pivot row rowmap[p] (or the unit bit).
__global__ void __launch_bounds__(256)
k_null_packed(const uint64_t *__restrict__ Q, size_t ldq, const int *__restrict__ rowmap,
              const int *__restrict__ freecols, size_t n, size_t f, uint64_t *__restrict__ Z,
              size_t ldz) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int rm = rowmap[p];
    const size_t fw = (f + 63) / 64;
    for (size_t jw = blockIdx.y; jw < ldz; jw += gridDim.y) {
        uint64_t word = 0;
        if (jw < fw) {
            if (rm >= 0) {
                const uint64_t *row = Q + (size_t)rm * ldq;
                const size_t jn = f - jw * 64 < 64 ? f - jw * 64 : 64;
                for (size_t b = 0; b < jn; ++b) {
                    const int fc = freecols[jw * 64 + b];
                    word |= ((row[fc >> 6] >> (fc & 63)) & 1ull) << b;
                }
            } else {
                const size_t fi = (size_t)(-rm - 1);
                if (fi / 64 == jw) word = 1ull << (fi & 63);
            }
        }
        Z[p * ldz + jw] = word;
    }
}

}  // namespace

mela_status eliminate(mela_ctx *ctx, const uint64_t *A, size_t lda, size_t m, size_t n,
                      int variant, bool rref, bool want_pivots, ElimResult *res,
                      uint64_t **Q_out, size_t *ldq_out) {
    hipStream_t s = ctx->stream;
    res->rank = 0;
    res->pivots.clear();
    if (Q_out) *Q_out = nullptr;
    const size_t wc = words64(n);
    const size_t rmax = m < n ? m : n;
    if (ldq_out) *ldq_out = wc ? wc : 1;
    if (m == 0 || n == 0) return MELA_OK;
    if (m > (size_t)INT32_MAX || wc > (size_t)INT32_MAX)
        return set_error(MELA_ERR_INVALID, "matrix dimension exceeds 2^31-1");

    const size_t ld = wc;
    void *pM = nullptr, *pQ = nullptr, *pMeta = nullptr, *pAct = nullptr;
    mela_status st;
    if ((st = scratch_get(ctx, SLOT_ELIM_M, m * ld * 8, &pM)) != MELA_OK) return st;
    if ((st = scratch_get(ctx, SLOT_ELIM_Q, rmax * ld * 8, &pQ)) != MELA_OK) return st;
    if ((st = scratch_get(ctx, SLOT_ELIM_META, sizeof(ElimMeta) + (rmax + 1) * sizeof(int), &pMeta)) !=
        MELA_OK)
        return st;
    if ((st = scratch_get(ctx, SLOT_ELIM_ACT, m * sizeof(uint32_t), &pAct)) != MELA_OK) return st;
    uint64_t *Mw = (uint64_t *)pM, *Q = (uint64_t *)pQ;
    ElimMeta *meta = (ElimMeta *)pMeta;
    int *pivcols = (int *)((char *)pMeta + sizeof(ElimMeta));
    uint32_t *act = (uint32_t *)pAct;

    const size_t tot = m * wc;
    hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, A, lda,
                       Mw, ld, m, wc);
    hipLaunchKernelGGL(k_elim_init, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, meta, act,
                       (int)m);
    MELA_HIP(hipGetLastError());

    const int cu = ctx->cu_count > 0 ? ctx->cu_count : 256;
    for (size_t g = 0; g < wc; ++g) {
        hipLaunchKernelGGL(k_panel_select, dim3(1), dim3(PS_THREADS), 0, s, Mw, ld, act, meta,
                           pivcols, (int)g, (int)rmax);
        dim3 gb((unsigned)((wc + 255) / 256), 64);
        hipLaunchKernelGGL(k_build_pivots, gb, dim3(256), 0, s, Mw, ld, Q, ld, meta, (int)g,
                           (int)wc);
        const size_t rem = wc - g - 1;
        if (rem > 0) {
            const unsigned tx = (unsigned)((rem + UP_TILE - 1) / UP_TILE);
            // rows are strided over gridDim.y; size it to fill the chip ~2x
            const size_t rows_hint = m + (rref ? rmax : 0);
            unsigned ty = (unsigned)((2 * cu + tx - 1) / tx);
            const unsigned ty_max = (unsigned)((rows_hint + 3) / 4);
            if (ty > ty_max) ty = ty_max;
            if (ty < 1) ty = 1;
            if (variant == MELA_ELIM_GE)
                hipLaunchKernelGGL(k_update_ge, dim3(tx, ty), dim3(256), 0, s, Mw, ld, act, Q, ld,
                                   meta, (int)g, (int)wc, rref ? 1 : 0);
            else
                hipLaunchKernelGGL(k_update_m4ri, dim3(tx, ty), dim3(256), 0, s, Mw, ld, act, Q,
                                   ld, meta, (int)g, (int)wc, rref ? 1 : 0);
        }
        if (rref)
            hipLaunchKernelGGL(k_update_panelword, dim3((unsigned)((rmax + 255) / 256)), dim3(256),
                               0, s, Q, ld, meta, (int)g);
        MELA_HIP(hipGetLastError());
    }
    int r_dev = 0;
    MELA_HIP(hipMemcpyAsync(&r_dev, &meta->r, sizeof(int), hipMemcpyDeviceToHost, s));
    MELA_HIP(hipStreamSynchronize(s));
    res->rank = (size_t)r_dev;
    if (want_pivots && r_dev > 0) {
        res->pivots.resize((size_t)r_dev);
        MELA_HIP(hipMemcpyAsync(res->pivots.data(), pivcols, (size_t)r_dev * sizeof(int),
                                hipMemcpyDeviceToHost, s));
        MELA_HIP(hipStreamSynchronize(s));
    }
    if (Q_out) *Q_out = Q;
    return MELA_OK;
}

mela_status launch_null_logical(hipStream_t s, const uint64_t *Q, size_t ldq, const int *rowmap,
                                const int *freecols, size_t n, size_t f, uint8_t *Z) {
    if (n == 0 || f == 0) return MELA_OK;
    const unsigned gy = (unsigned)(f < 65535 ? f : 65535);
    hipLaunchKernelGGL(k_null_logical, dim3((unsigned)((n + 255) / 256), gy), dim3(256), 0, s, Q,
                       ldq, rowmap, freecols, n, f, Z);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

mela_status launch_null_packed(hipStream_t s, const uint64_t *Q, size_t ldq, const int *rowmap,
                               const int *freecols, size_t n, size_t f, uint64_t *Z, size_t ldz) {
    if (n == 0 || ldz == 0) return MELA_OK;
    const unsigned gy = (unsigned)(ldz < 65535 ? ldz : 65535);
    hipLaunchKernelGGL(k_null_packed, dim3((unsigned)((n + 255) / 256), gy), dim3(256), 0, s, Q,
                       ldq, rowmap, freecols, n, f, Z, ldz);
    MELA_HIP(hipGetLastError());
    return MELA_OK;
}

}  // namespace mela
