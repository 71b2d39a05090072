/*
 * mela_oracle.c -- CPU restatement of the six HCPhy/mex_linear_algebra C sources.
 *
 * TEST INFRASTRUCTURE ONLY (see mela_oracle.h).  Written from the reference's
 * observable algorithm, not copied: each function names the reference lines
 * whose loop order, blocking, pivot rule and OpenMP placement it follows, so
 * that (a) its outputs are those of the reference and (b) its timings stand in
 * for the reference's CPU path on a machine where MATLAB is absent.
 *
 * Build flags follow compile_mex.m:47-56 (-O3 -mavx512f -mavx512bw -mavx512dq
 * -fopenmp, or -O3 -mavx2 -fopenmp without AVX-512F).  The XOR loops are
 * plain C; gcc vectorises them to the same 512-/256-bit XORs the reference
 * writes with intrinsics (matmul_m4ri.c:42-75).
 */
#include "mela_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define W64(n) (((n) + 63) >> 6)

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ((int)x) & 1 as the reference evaluates it (matmul_gf2.c:66,139;
 * rank_gf2.c:154; rank_m4ri.c:312; null_gf2.c:173).  For |x| < 2^31 this is
 * truncation toward zero then the two's-complement low bit.  The conversion
 * of NaN, +-Inf and values outside int range is undefined in C; on x86-64 the
 * cvttsd2si/vcvttpd2dq instructions gcc emits return the "integer indefinite"
 * 0x80000000, whose low bit is 0.  This restatement fixes that x86 result. */
int orc_parity_double(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0))
        return 0;
    return ((int)x) & 1;
}

static inline int elem_bit(const void *in, int kind, size_t idx)
{
    if (kind == ORC_DOUBLE)
        return orc_parity_double(((const double *)in)[idx]);
    return ((const uint8_t *)in)[idx] != 0;
}

/* Parity packers: matmul_gf2.c:42-69 (row-outer) and rank_m4ri.c:308-334 /
 * null_gf2.c:169-187 (column-outer).  The packed result does not depend on the
 * loop order; column-outer is used because it reads the column-major source
 * sequentially. */
void orc_pack(const void *in, int kind, size_t rows, size_t cols,
              uint64_t *out, size_t ld)
{
    memset(out, 0, rows * ld * sizeof(uint64_t));
    for (size_t j = 0; j < cols; ++j)
        for (size_t i = 0; i < rows; ++i)
            if (elem_bit(in, kind, i + j * rows))
                out[i * ld + (j >> 6)] |= (uint64_t)1 << (j & 63);
}

/* Unpack to column-major logical: matmul_gf2.c:193-205. */
void orc_unpack(const uint64_t *in, size_t ld, size_t rows, size_t cols,
                uint8_t *out)
{
    memset(out, 0, rows * cols);
    for (size_t i = 0; i < rows; ++i)
        for (size_t j = 0; j < cols; ++j)
            if ((in[i * ld + (j >> 6)] >> (j & 63)) & 1)
                out[i + j * rows] = 1;
}

static inline void xor_words(uint64_t *restrict d, const uint64_t *restrict s,
                             size_t nw)
{
    for (size_t w = 0; w < nw; ++w)
        d[w] ^= s[w];
}

#define BLOCK_ROWS 256

/* ------------------------------------------------------------------ matmul */

/* mela_matmul_gf2 core, matmul_gf2.c:90-208: pack B by rows once, then an
 * OpenMP loop over 256-row blocks of C; inside a block iterate k then i and
 * XOR packed row B_k into C_i whenever A(i,k) is odd; unpack the block. */
void orc_matmul_gf2(const void *A, int ka, size_t m, size_t k,
                    const void *B, int kb, size_t n, uint8_t *C)
{
    size_t wb = W64(n);
    uint64_t *Bp = (uint64_t *)calloc(k * wb + 1, sizeof(uint64_t));
    orc_pack(B, kb, k, n, Bp, wb);
    memset(C, 0, m * n);
#pragma omp parallel for schedule(static)
    for (size_t ib = 0; ib < m; ib += BLOCK_ROWS) {
        uint64_t *Cb = (uint64_t *)calloc((size_t)BLOCK_ROWS * wb + 1, sizeof(uint64_t));
        size_t h = (ib + BLOCK_ROWS < m) ? BLOCK_ROWS : m - ib;
        for (size_t kk = 0; kk < k; ++kk) {
            const uint64_t *Bk = Bp + kk * wb;
            for (size_t i = 0; i < h; ++i)
                if (elem_bit(A, ka, (ib + i) + kk * m))
                    xor_words(Cb + i * wb, Bk, wb);
        }
        for (size_t i = 0; i < h; ++i)
            for (size_t j = 0; j < n; ++j)
                if ((Cb[i * wb + (j >> 6)] >> (j & 63)) & 1)
                    C[(ib + i) + j * m] = 1;
        free(Cb);
    }
    free(Bp);
}

/* T[0]=0, T[j + 2^b] = T[j] ^ rows[b] (doubling), matmul_m4ri.c:113-137. */
static void m4rm_table(uint64_t *T, const uint64_t *rows, size_t ldr, int kk,
                       size_t nw)
{
    memset(T, 0, nw * sizeof(uint64_t));
    size_t size = 1;
    for (int b = 0; b < kk; ++b) {
        const uint64_t *rb = rows + (size_t)b * ldr;
        for (size_t j = 0; j < size; ++j) {
            uint64_t *d = T + (size + j) * nw;
            memcpy(d, T + j * nw, nw * sizeof(uint64_t));
            xor_words(d, rb, nw);
        }
        size <<= 1;
    }
}

/* mela_matmul_m4ri core, matmul_m4ri.c:158-248: per 256-row block and per
 * 8-column strip of A, build the 256-entry table of B rows, gather the 8 A
 * bits of each row into an index (bit b <-> column k_base+b), XOR T[idx]. */
void orc_matmul_m4ri(const void *A, int ka, size_t m, size_t k,
                     const void *B, int kb, size_t n, uint8_t *C)
{
    size_t wb = W64(n);
    uint64_t *Bp = (uint64_t *)calloc(k * wb + 1, sizeof(uint64_t));
    orc_pack(B, kb, k, n, Bp, wb);
    memset(C, 0, m * n);
#pragma omp parallel for schedule(static)
    for (size_t ib = 0; ib < m; ib += BLOCK_ROWS) {
        uint64_t *Cb = (uint64_t *)calloc((size_t)BLOCK_ROWS * wb + 1, sizeof(uint64_t));
        uint64_t *T = (uint64_t *)calloc((size_t)256 * wb + 1, sizeof(uint64_t));
        size_t h = (ib + BLOCK_ROWS < m) ? BLOCK_ROWS : m - ib;
        for (size_t kb0 = 0; kb0 < k; kb0 += 8) {
            int kc = (kb0 + 8 <= k) ? 8 : (int)(k - kb0);
            m4rm_table(T, Bp + kb0 * wb, wb, kc, wb);
            for (size_t i = 0; i < h; ++i) {
                unsigned idx = 0;
                for (int b = 0; b < kc; ++b)
                    if (elem_bit(A, ka, (ib + i) + (kb0 + b) * m))
                        idx |= 1u << b;
                if (idx)
                    xor_words(Cb + i * wb, T + idx * wb, wb);
            }
        }
        for (size_t i = 0; i < h; ++i)
            for (size_t j = 0; j < n; ++j)
                if ((Cb[i * wb + (j >> 6)] >> (j & 63)) & 1)
                    C[(ib + i) + j * m] = 1;
        free(T);
        free(Cb);
    }
    free(Bp);
}

/* Kernel-only form of matmul_gf2.c:110-190 on a packed A (bit test instead of
 * reading the MATLAB array); C is written packed. */
void orc_matmul_gf2_packed(const uint64_t *A, size_t lda, size_t m, size_t k,
                           const uint64_t *B, size_t ldb, size_t n,
                           uint64_t *C, size_t ldc)
{
    size_t wb = W64(n);
#pragma omp parallel for schedule(static)
    for (size_t ib = 0; ib < m; ib += BLOCK_ROWS) {
        size_t h = (ib + BLOCK_ROWS < m) ? BLOCK_ROWS : m - ib;
        for (size_t i = 0; i < h; ++i)
            memset(C + (ib + i) * ldc, 0, ldc * sizeof(uint64_t));
        for (size_t kk = 0; kk < k; ++kk) {
            const uint64_t *Bk = B + kk * ldb;
            for (size_t i = 0; i < h; ++i)
                if ((A[(ib + i) * lda + (kk >> 6)] >> (kk & 63)) & 1)
                    xor_words(C + (ib + i) * ldc, Bk, wb);
        }
    }
}

/* Kernel-only form of matmul_m4ri.c:178-229 on packed operands. */
void orc_matmul_m4ri_packed(const uint64_t *A, size_t lda, size_t m, size_t k,
                            const uint64_t *B, size_t ldb, size_t n,
                            uint64_t *C, size_t ldc)
{
    size_t wb = W64(n);
#pragma omp parallel for schedule(static)
    for (size_t ib = 0; ib < m; ib += BLOCK_ROWS) {
        uint64_t *T = (uint64_t *)calloc((size_t)256 * wb + 1, sizeof(uint64_t));
        size_t h = (ib + BLOCK_ROWS < m) ? BLOCK_ROWS : m - ib;
        for (size_t i = 0; i < h; ++i)
            memset(C + (ib + i) * ldc, 0, ldc * sizeof(uint64_t));
        for (size_t kb0 = 0; kb0 < k; kb0 += 8) {
            int kc = (kb0 + 8 <= k) ? 8 : (int)(k - kb0);
            m4rm_table(T, B + kb0 * ldb, ldb, kc, wb);
            for (size_t i = 0; i < h; ++i) {
                unsigned idx = (unsigned)((A[(ib + i) * lda + (kb0 >> 6)] >> (kb0 & 63)) & 0xff);
                if (kc < 8)
                    idx &= (1u << kc) - 1u;
                if (idx)
                    xor_words(C + (ib + i) * ldc, T + idx * wb, wb);
            }
        }
        free(T);
    }
}

/* -------------------------------------------------------------------- rank */

/* bitrank_packed, rank_gf2.c:27-121: for each column, the pivot is the first
 * row at or below the current row with the bit set (serial scan), swap it up,
 * then an OpenMP loop XORs the whole pivot row into every lower row with the
 * bit set. */
double orc_rank_gf2_packed(uint64_t *M, size_t m, size_t n, size_t ld)
{
    double rk = 0.0;
    size_t i = 0, col = 0;
    while (i < m && col < n) {
        size_t wi = col >> 6;
        uint64_t mask = (uint64_t)1 << (col & 63);
        size_t p = m;
        for (size_t j = i; j < m; ++j)
            if (M[j * ld + wi] & mask) { p = j; break; }
        if (p == m) { ++col; continue; }
        if (p != i)
            for (size_t w = 0; w < ld; ++w) {
                uint64_t t = M[i * ld + w];
                M[i * ld + w] = M[p * ld + w];
                M[p * ld + w] = t;
            }
        rk += 1.0;
        const uint64_t *pr = M + i * ld;
#pragma omp parallel for schedule(static)
        for (size_t j = i + 1; j < m; ++j)
            if (M[j * ld + wi] & mask)
                xor_words(M + j * ld, pr, ld);
        ++i;
        ++col;
    }
    return rk;
}

#define RK_K 10

/* rank_m4ri, rank_m4ri.c:94-279 with K=10: per block of up to 10 columns,
 * pivots are searched in a window of rows [r, limit) that grows by 10 rows at
 * a time (new rows are first cleaned against the block's pivots); each pivot
 * is swapped to r+p and cleared from the window; the block's pivots are then
 * reduced against each other (Gauss-Jordan), a 2^p table is built and an
 * OpenMP loop clears the block's pivot columns from every row at or below
 * the window by one table XOR per row. */
double orc_rank_m4ri_packed(uint64_t *M, size_t m, size_t n, size_t ld)
{
    double rk = 0.0;
    size_t r = 0, c = 0;
    uint64_t *T = (uint64_t *)malloc(((size_t)1 << RK_K) * ld * sizeof(uint64_t) + 8);
    uint64_t *piv[RK_K];
    size_t pidx[RK_K];
    size_t pw[RK_K];
    uint64_t pm[RK_K];
    while (r < m && c < n) {
        int kc = (c + RK_K <= n) ? RK_K : (int)(n - c);
        if (kc <= 0)
            break;
        int pc = 0;
        size_t limit = r + RK_K;
        if (limit > m)
            limit = m;
        for (int kk = 0; kk < kc; ++kk) {
            if (r + pc >= m)
                break;
            size_t col = c + (size_t)kk;
            size_t wi = col >> 6;
            uint64_t mask = (uint64_t)1 << (col & 63);
            size_t prow = r + pc;
            int found = 0;
            for (size_t i = r + pc; i < limit; ++i)
                if (M[i * ld + wi] & mask) { prow = i; found = 1; break; }
            while (!found && limit < m) {
                size_t old = limit;
                limit = (limit + RK_K <= m) ? limit + RK_K : m;
                for (int p = 0; p < pc; ++p) {
                    size_t pwi = pidx[p] >> 6;
                    uint64_t pmask = (uint64_t)1 << (pidx[p] & 63);
                    for (size_t j = old; j < limit; ++j)
                        if (M[j * ld + pwi] & pmask)
                            xor_words(M + j * ld, piv[p], ld);
                }
                for (size_t i = old; i < limit; ++i)
                    if (M[i * ld + wi] & mask) { prow = i; found = 1; break; }
            }
            if (!found)
                continue;
            if (prow != r + pc)
                for (size_t w = 0; w < ld; ++w) {
                    uint64_t t = M[(r + pc) * ld + w];
                    M[(r + pc) * ld + w] = M[prow * ld + w];
                    M[prow * ld + w] = t;
                }
            uint64_t *pp = M + (r + pc) * ld;
            for (size_t j = r + pc + 1; j < limit; ++j)
                if (M[j * ld + wi] & mask)
                    xor_words(M + j * ld, pp, ld);
            piv[pc] = pp;
            pidx[pc] = col;
            ++pc;
        }
        if (pc > 0) {
            for (int a = 0; a < pc; ++a) {
                size_t wi = pidx[a] >> 6;
                uint64_t mask = (uint64_t)1 << (pidx[a] & 63);
                for (int b = 0; b < a; ++b)
                    if (piv[b][wi] & mask)
                        xor_words(piv[b], piv[a], ld);
            }
            for (int a = 0; a < pc; ++a) {
                pw[a] = pidx[a] >> 6;
                pm[a] = (uint64_t)1 << (pidx[a] & 63);
            }
            /* table of all 2^pc combinations (rank_m4ri.c:65-88) */
            memset(T, 0, ld * sizeof(uint64_t));
            size_t size = 1;
            for (int a = 0; a < pc; ++a) {
                for (size_t j = 0; j < size; ++j) {
                    uint64_t *d = T + (size + j) * ld;
                    memcpy(d, T + j * ld, ld * sizeof(uint64_t));
                    xor_words(d, piv[a], ld);
                }
                size <<= 1;
            }
#pragma omp parallel for schedule(static)
            for (size_t j = limit; j < m; ++j) {
                uint64_t *row = M + j * ld;
                unsigned idx = 0;
                for (int a = 0; a < pc; ++a)
                    if (row[pw[a]] & pm[a])
                        idx |= 1u << a;
                if (idx)
                    xor_words(row, T + (size_t)idx * ld, ld);
            }
            rk += (double)pc;
            r += (size_t)pc;
        }
        c += (size_t)kc;
    }
    free(T);
    return rk;
}

/* -------------------------------------------------------------------- rref */

/* gf2_rref_standard, null_gf2.c:88-131: pivot = first row at or below the
 * pivot row with the bit; swap; OpenMP loop clears the column from every
 * other row (full RREF). */
long orc_rref_gf2_packed(uint64_t *M, size_t m, size_t n, size_t ld, int *pivots)
{
    size_t prow = 0;
    long cnt = 0;
    for (size_t c = 0; c < n && prow < m; ++c) {
        size_t wi = c >> 6;
        uint64_t mask = (uint64_t)1 << (c & 63);
        size_t p = m;
        for (size_t r = prow; r < m; ++r)
            if (M[r * ld + wi] & mask) { p = r; break; }
        if (p == m)
            continue;
        if (p != prow)
            for (size_t w = 0; w < ld; ++w) {
                uint64_t t = M[p * ld + w];
                M[p * ld + w] = M[prow * ld + w];
                M[prow * ld + w] = t;
            }
        pivots[cnt++] = (int)c;
        const uint64_t *pr = M + prow * ld;
#pragma omp parallel for schedule(static)
        for (size_t r = 0; r < m; ++r)
            if (r != prow && (M[r * ld + wi] & mask))
                xor_words(M + r * ld, pr, ld);
        ++prow;
    }
    return cnt;
}

#define NL_K 8

/* gf2_rref_m4ri, null_m4ri.c:92-240: collect up to 8 pivots scanning columns
 * from col_base (pivot = first row at or below pivot_row+k with the bit, swap,
 * then serial eager elimination of that column from every lower row); build
 * the 2^k table by lowest-set-bit decomposition; an OpenMP loop XORs
 * T[idx(row)] into every row outside the block's pivot rows. */
long orc_rref_m4ri_packed(uint64_t *M, size_t m, size_t n, size_t ld, int *pivots)
{
    size_t prow = 0;
    long cnt = 0;
    size_t cb = 0;
    while (cb < n && prow < m) {
        int kp = 0;
        int cp[NL_K];
        size_t cr[NL_K];
        size_t c = cb;
        while (c < n && kp < NL_K && prow + kp < m) {
            size_t start = prow + kp;
            size_t wi = c >> 6;
            uint64_t mask = (uint64_t)1 << (c & 63);
            size_t p = m;
            for (size_t r = start; r < m; ++r)
                if (M[r * ld + wi] & mask) { p = r; break; }
            if (p != m) {
                if (p != start)
                    for (size_t w = 0; w < ld; ++w) {
                        uint64_t t = M[p * ld + w];
                        M[p * ld + w] = M[start * ld + w];
                        M[start * ld + w] = t;
                    }
                cp[kp] = (int)c;
                cr[kp] = start;
                pivots[cnt++] = (int)c;
                const uint64_t *cur = M + start * ld;
                for (size_t r = start + 1; r < m; ++r)
                    if (M[r * ld + wi] & mask)
                        xor_words(M + r * ld, cur, ld);
                ++kp;
            }
            ++c;
        }
        if (kp == 0) {
            if (c >= n)
                break;
            cb = c;
            continue;
        }
        size_t ne = (size_t)1 << kp;
        uint64_t *T = (uint64_t *)calloc(ne * ld + 1, sizeof(uint64_t));
        for (size_t i = 1; i < ne; ++i) {
            int b = 0;
            while (!((i >> b) & 1))
                ++b;
            size_t prev = i ^ ((size_t)1 << b);
            memcpy(T + i * ld, T + prev * ld, ld * sizeof(uint64_t));
            xor_words(T + i * ld, M + cr[b] * ld, ld);
        }
#pragma omp parallel for schedule(static)
        for (size_t r = 0; r < m; ++r) {
            if (r >= prow && r < prow + (size_t)kp)
                continue;
            const uint64_t *row = M + r * ld;
            unsigned idx = 0;
            for (int j = 0; j < kp; ++j)
                if ((row[(size_t)cp[j] >> 6] >> (cp[j] & 63)) & 1)
                    idx |= 1u << j;
            if (idx)
                xor_words(M + r * ld, T + (size_t)idx * ld, ld);
        }
        free(T);
        prow += (size_t)kp;
        cb = c;
    }
    return cnt;
}

/* Back-substitution, null_gf2.c:193-274 (identical in null_m4ri.c:302-401):
 * free columns in ascending order; for each, v = e_f, then for pivots from
 * the last to the first x[c] = parity(v AND row_pi without its pivot bit);
 * column fv of the n x (n-r) logical output is v. */
void orc_null_backsub(const uint64_t *M, size_t n, size_t ld,
                      const int *pivots, long r, uint8_t *Z)
{
    size_t wc = W64(n);
    uint8_t *isp = (uint8_t *)calloc(n + 1, 1);
    for (long i = 0; i < r; ++i)
        isp[pivots[i]] = 1;
    uint64_t *v = (uint64_t *)calloc(wc + 1, sizeof(uint64_t));
    uint64_t *sv = (uint64_t *)calloc(wc + 1, sizeof(uint64_t));
    size_t fi = 0;
    size_t nf = n - (size_t)r;
    memset(Z, 0, n * nf);
    for (size_t f = 0; f < n; ++f) {
        if (isp[f])
            continue;
        memset(v, 0, wc * sizeof(uint64_t));
        v[f >> 6] |= (uint64_t)1 << (f & 63);
        for (long pi = r - 1; pi >= 0; --pi) {
            int c = pivots[pi];
            memcpy(sv, M + (size_t)pi * ld, wc * sizeof(uint64_t));
            sv[c >> 6] &= ~((uint64_t)1 << (c & 63));
            uint64_t acc = 0;
            for (size_t w = 0; w < wc; ++w)
                acc ^= v[w] & sv[w];
            int par = __builtin_parityll(acc);
            if (par)
                v[c >> 6] ^= (uint64_t)1 << (c & 63);
            else
                v[c >> 6] &= ~((uint64_t)1 << (c & 63));
        }
        for (size_t j = 0; j < n; ++j)
            Z[j + fi * n] = (uint8_t)((v[j >> 6] >> (j & 63)) & 1);
        ++fi;
    }
    free(sv);
    free(v);
    free(isp);
}

double orc_rank_gf2(const void *A, int ka, size_t m, size_t n)
{
    size_t wc = W64(n);
    uint64_t *M = (uint64_t *)calloc(m * wc + 1, sizeof(uint64_t));
    orc_pack(A, ka, m, n, M, wc);
    double r = orc_rank_gf2_packed(M, m, n, wc);
    free(M);
    return r;
}

/* rank_m4ri gateway pads rows to a multiple of 8 words (rank_m4ri.c:297-298) */
double orc_rank_m4ri(const void *A, int ka, size_t m, size_t n)
{
    size_t ld = (W64(n) + 7) & ~(size_t)7;
    uint64_t *M = (uint64_t *)calloc(m * ld + 1, sizeof(uint64_t));
    orc_pack(A, ka, m, n, M, ld);
    double r = orc_rank_m4ri_packed(M, m, n, ld);
    free(M);
    return r;
}

static long null_common(const void *H, int kh, size_t m, size_t n, uint8_t *Z,
                        int m4ri)
{
    size_t wc = W64(n);
    uint64_t *M = (uint64_t *)calloc(m * wc + 1, sizeof(uint64_t));
    orc_pack(H, kh, m, n, M, wc);
    size_t mn = m < n ? m : n;
    int *piv = (int *)calloc(mn + 1, sizeof(int));
    long r = m4ri ? orc_rref_m4ri_packed(M, m, n, wc, piv)
                  : orc_rref_gf2_packed(M, m, n, wc, piv);
    orc_null_backsub(M, n, wc, piv, r, Z);
    free(piv);
    free(M);
    return (long)n - r;
}

long orc_null_gf2(const void *H, int kh, size_t m, size_t n, uint8_t *Z)
{
    return null_common(H, kh, m, n, Z, 0);
}

long orc_null_m4ri(const void *H, int kh, size_t m, size_t n, uint8_t *Z)
{
    return null_common(H, kh, m, n, Z, 1);
}

/* ---------------------------------------------------------------- generator */

uint64_t orc_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* word (i,w) = splitmix64(splitmix64(k1 + i*ceil(cols/64) + w) ^ k2) with
 * k1 = splitmix64(seed), k2 = splitmix64(seed ^ 0xD1B54A32D192ED03); bits past
 * the last column cleared.  Two rounds: a single splitmix64 round on a
 * counter leaves GF(2)-linear dependences between bit columns (a 12000 x 11968
 * block had rank 11855), which would make "random" matrices rank-deficient.
 * The device generator computes the same words. */
void orc_gen_bernoulli_packed(uint64_t seed, size_t rows, size_t cols,
                              uint64_t *out, size_t ld)
{
    size_t wc = W64(cols);
    uint64_t k1 = orc_splitmix64(seed);
    uint64_t k2 = orc_splitmix64(seed ^ 0xD1B54A32D192ED03ULL);
    uint64_t last = (cols & 63) ? (((uint64_t)1 << (cols & 63)) - 1) : ~(uint64_t)0;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < rows; ++i) {
        for (size_t w = 0; w < ld; ++w) {
            uint64_t v = 0;
            if (w < wc) {
                v = orc_splitmix64(orc_splitmix64(k1 + (uint64_t)(i * wc + w)) ^ k2);
                if (w == wc - 1)
                    v &= last;
            }
            out[i * ld + w] = v;
        }
    }
}
