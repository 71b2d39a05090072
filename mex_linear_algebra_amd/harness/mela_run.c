/*
 * mela_run -- C harness for libmela without MATLAB or Python.
 *
 *   mela_run selftest             six entry points on seeded inputs, checked
 *                                 against naive C (A*B mod 2, row reduction)
 *   mela_run matmul M K N [algo]  packed device product, ms and Tbitop/s
 *   mela_run rank M N [ge|m4ri]   packed device rank, ms
 *   mela_run null M N [ge|m4ri]   packed device null basis, ms and nullity
 *
 * Inputs are Bernoulli(1/2) from the library's seeded generator (seed 1, 2).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mela.h"

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static void rand_logical(uint8_t *a, size_t n, uint64_t seed) {
    for (size_t i = 0; i < n; ++i) a[i] = (uint8_t)(sm64(&seed) & 1);
}

/* naive column-major rank by row reduction on a byte copy */
static size_t naive_rank(const uint8_t *A, size_t m, size_t n, int *pivcols) {
    uint8_t *M = malloc(m * n + 1);
    memcpy(M, A, m * n);
    size_t row = 0;
    for (size_t c = 0; c < n && row < m; ++c) {
        size_t p = m;
        for (size_t r = row; r < m; ++r)
            if (M[r + c * m]) { p = r; break; }
        if (p == m) continue;
        for (size_t j = 0; j < n; ++j) {
            uint8_t t = M[row + j * m];
            M[row + j * m] = M[p + j * m];
            M[p + j * m] = t;
        }
        for (size_t r = 0; r < m; ++r)
            if (r != row && M[r + c * m])
                for (size_t j = 0; j < n; ++j) M[r + j * m] ^= M[row + j * m];
        if (pivcols) pivcols[row] = (int)c;
        ++row;
    }
    free(M);
    return row;
}

typedef struct { uint8_t *buf; size_t rows, cols; } out_t;
static uint8_t *alloc_cb(void *user, size_t rows, size_t cols) {
    out_t *o = (out_t *)user;
    o->rows = rows;
    o->cols = cols;
    o->buf = calloc(rows * cols + 1, 1);
    return o->buf;
}

#define CHECK(cond, ...)                         \
    do {                                         \
        if (!(cond)) {                           \
            fprintf(stderr, "FAIL: " __VA_ARGS__); \
            fprintf(stderr, "\n");               \
            return 1;                            \
        }                                        \
    } while (0)

static int selftest(void) {
    const size_t shapes[][3] = {{37, 23, 31}, {64, 64, 64}, {130, 70, 257}, {5, 7, 3}};
    for (size_t s = 0; s < sizeof shapes / sizeof shapes[0]; ++s) {
        size_t m = shapes[s][0], k = shapes[s][1], n = shapes[s][2];
        uint8_t *A = malloc(m * k), *B = malloc(k * n), *C = malloc(m * n), *W = calloc(m * n, 1);
        rand_logical(A, m * k, 10 + s);
        rand_logical(B, k * n, 20 + s);
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) {
                unsigned acc = 0;
                for (size_t q = 0; q < k; ++q) acc ^= A[i + q * m] & B[q + j * k];
                W[i + j * m] = (uint8_t)acc;
            }
        for (int v = 0; v < 2; ++v) {
            mela_status st = v ? mela_matmul_m4ri(NULL, A, MELA_LOGICAL, m, k, B, MELA_LOGICAL, k, n, C)
                               : mela_matmul_gf2(NULL, A, MELA_LOGICAL, m, k, B, MELA_LOGICAL, k, n, C);
            CHECK(st == MELA_OK, "matmul status %d: %s", st, mela_last_error());
            CHECK(!memcmp(C, W, m * n), "matmul mismatch %zux%zux%zu v%d", m, k, n, v);
        }
        /* rank and null of A */
        int *piv = malloc(sizeof(int) * (m < k ? m : k) + 4);
        size_t r = naive_rank(A, m, k, piv);
        double r1 = -1, r2 = -1;
        CHECK(mela_rank_gf2(NULL, A, MELA_LOGICAL, m, k, &r1) == MELA_OK, "rank_gf2");
        CHECK(mela_rank_m4ri(NULL, A, MELA_LOGICAL, m, k, &r2) == MELA_OK, "rank_m4ri");
        CHECK(r1 == (double)r && r2 == (double)r, "rank %g %g vs %zu", r1, r2, r);
        for (int v = 0; v < 2; ++v) {
            out_t o = {0};
            size_t f = 0;
            mela_status st = v ? mela_null_m4ri(NULL, A, MELA_LOGICAL, m, k, alloc_cb, &o, &f)
                               : mela_null_gf2(NULL, A, MELA_LOGICAL, m, k, alloc_cb, &o, &f);
            CHECK(st == MELA_OK, "null status %d", st);
            CHECK(f == k - r && o.rows == k && o.cols == f, "null shape");
            /* A*Z = 0 */
            for (size_t i = 0; i < m; ++i)
                for (size_t j = 0; j < f; ++j) {
                    unsigned acc = 0;
                    for (size_t q = 0; q < k; ++q) acc ^= A[i + q * m] & o.buf[q + j * k];
                    CHECK(acc == 0, "A*Z != 0");
                }
            free(o.buf);
        }
        free(piv);
        free(A);
        free(B);
        free(C);
        free(W);
    }
    printf("selftest ok\n");
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: mela_run selftest | matmul M K N [algo] | rank M N [ge|m4ri] | null M N [ge|m4ri]\n");
        return 2;
    }
    if (!strcmp(argv[1], "selftest")) return selftest();
    mela_ctx *ctx = NULL;
    if (mela_ctx_create(0, &ctx) != MELA_OK) {
        fprintf(stderr, "%s\n", mela_last_error());
        return 1;
    }
    int rc = 0;
    if (!strcmp(argv[1], "matmul") && argc >= 5) {
        size_t m = strtoull(argv[2], 0, 10), k = strtoull(argv[3], 0, 10), n = strtoull(argv[4], 0, 10);
        int algo = argc > 5 ? atoi(argv[5]) : MELA_ALGO_AUTO;
        mela_bitmat A, B, C;
        mela_bitmat_alloc(ctx, m, k, &A);
        mela_bitmat_alloc(ctx, k, n, &B);
        mela_bitmat_alloc(ctx, m, n, &C);
        mela_gen_bernoulli(ctx, 1, &A);
        mela_gen_bernoulli(ctx, 2, &B);
        mela_matmul_packed(ctx, &A, &B, &C, algo);
        mela_ctx_sync(ctx);
        double t0 = now_ms();
        const int it = 5;
        for (int i = 0; i < it; ++i) mela_matmul_packed(ctx, &A, &B, &C, algo);
        mela_ctx_sync(ctx);
        double ms = (now_ms() - t0) / it;
        printf("{\"op\":\"matmul\",\"m\":%zu,\"k\":%zu,\"n\":%zu,\"ms\":%.4f,\"tbitops\":%.2f}\n", m, k, n,
               ms, 2.0 * m * n * k / (ms * 1e-3) / 1e12);
        mela_bitmat_free(ctx, &A);
        mela_bitmat_free(ctx, &B);
        mela_bitmat_free(ctx, &C);
    } else if ((!strcmp(argv[1], "rank") || !strcmp(argv[1], "null")) && argc >= 4) {
        size_t m = strtoull(argv[2], 0, 10), n = strtoull(argv[3], 0, 10);
        int var = (argc > 4 && !strcmp(argv[4], "ge")) ? MELA_ELIM_GE : MELA_ELIM_M4RI;
        mela_bitmat A;
        mela_bitmat_alloc(ctx, m, n, &A);
        mela_gen_bernoulli(ctx, 1, &A);
        size_t r = 0;
        double t0 = now_ms();
        if (!strcmp(argv[1], "rank")) {
            if (mela_rank_packed(ctx, &A, var, &r) != MELA_OK) rc = 1;
        } else {
            mela_bitmat Z;
            if (mela_null_packed(ctx, &A, var, &Z, &r) != MELA_OK) rc = 1;
            else mela_bitmat_free(ctx, &Z);
        }
        double ms = now_ms() - t0;
        if (rc) fprintf(stderr, "%s\n", mela_last_error());
        printf("{\"op\":\"%s\",\"m\":%zu,\"n\":%zu,\"result\":%zu,\"ms\":%.3f}\n", argv[1], m, n, r, ms);
        mela_bitmat_free(ctx, &A);
    } else {
        fprintf(stderr, "bad arguments\n");
        rc = 2;
    }
    mela_ctx_destroy(ctx);
    return rc;
}
