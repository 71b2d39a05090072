/*
 * mela_oracle.h -- CPU restatement of the reference GF(2) kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline; it is never linked into, loaded by or called from the
 * product library (libmela.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it.
 *
 * Every function follows one reference function step for step (same loop
 * order, block sizes, pivot rule and OpenMP placement); the citation is on each
 * definition in mela_oracle.c.  Matrices are MATLAB column-major on the
 * "matlab" entry points and row-major bit-packed (bit j of row i at word j>>6,
 * bit j&63, LSB first) on the "packed" entry points.
 */
#ifndef MELA_ORACLE_H
#define MELA_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element parity exactly as ((int)x) & 1 evaluates with x86-64 cvttsd2si */
int orc_parity_double(double x);

/* input kinds for the matlab-layout entry points */
enum { ORC_LOGICAL = 0, ORC_DOUBLE = 1 };

void orc_pack(const void *in, int kind, size_t rows, size_t cols,
              uint64_t *out, size_t ld);
void orc_unpack(const uint64_t *in, size_t ld, size_t rows, size_t cols,
                uint8_t *out);

/* end-to-end restatements: MATLAB column-major in, logical/double out */
void orc_matmul_gf2(const void *A, int ka, size_t m, size_t k,
                    const void *B, int kb, size_t n, uint8_t *C);
void orc_matmul_m4ri(const void *A, int ka, size_t m, size_t k,
                     const void *B, int kb, size_t n, uint8_t *C);
double orc_rank_gf2(const void *A, int ka, size_t m, size_t n);
double orc_rank_m4ri(const void *A, int ka, size_t m, size_t n);
/* Z must hold n*n bytes (nullity <= n); returns the nullity */
long orc_null_gf2(const void *H, int kh, size_t m, size_t n, uint8_t *Z);
long orc_null_m4ri(const void *H, int kh, size_t m, size_t n, uint8_t *Z);

/* kernel-only restatements on packed rows (ld words per row) */
void orc_matmul_gf2_packed(const uint64_t *A, size_t lda, size_t m, size_t k,
                           const uint64_t *B, size_t ldb, size_t n,
                           uint64_t *C, size_t ldc);
void orc_matmul_m4ri_packed(const uint64_t *A, size_t lda, size_t m, size_t k,
                            const uint64_t *B, size_t ldb, size_t n,
                            uint64_t *C, size_t ldc);
double orc_rank_gf2_packed(uint64_t *M, size_t m, size_t n, size_t ld);
/* ld must be a multiple of 8 words (rank_m4ri pads rows to 64 bytes) */
double orc_rank_m4ri_packed(uint64_t *M, size_t m, size_t n, size_t ld);
long orc_rref_gf2_packed(uint64_t *M, size_t m, size_t n, size_t ld, int *pivots);
long orc_rref_m4ri_packed(uint64_t *M, size_t m, size_t n, size_t ld, int *pivots);
/* canonical null basis from an echelon form with pivot list; Z is n x (n-r) */
void orc_null_backsub(const uint64_t *M, size_t n, size_t ld,
                      const int *pivots, long r, uint8_t *Z);

/* seeded Bernoulli(1/2) packed generator shared with the device harness */
uint64_t orc_splitmix64(uint64_t x);
void orc_gen_bernoulli_packed(uint64_t seed, size_t rows, size_t cols,
                              uint64_t *out, size_t ld);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
