// The six reference entry points and the device-resident packed API.
//
// Each entry point restates the validation of the matching mexFunction (same
// order, same message text and identifiers), then runs pack -> kernels ->
// unpack on the device.  Host arrays are MATLAB column-major.
#include <cstring>
#include <vector>

#include "common.h"
#include "hostpack.h"

using namespace mela;

namespace {

const char *kSparseMsg = "Sparse matrices not supported. Convert to full using full().";

size_t elem_size(int cls) { return (cls & 0xff) == MELA_DOUBLE ? 8 : 1; }
bool is_supported(int cls) {
    const int c = cls & 0xff;
    return c == MELA_LOGICAL || c == MELA_DOUBLE;
}

bool mul_overflows(size_t a, size_t b, size_t *out) {
    if (a && b > SIZE_MAX / a) return true;
    *out = a * b;
    return false;
}

// ctx == NULL selects the default context (after validation has passed)
mela_status resolve(mela_ctx **ctx) {
    if (*ctx) return MELA_OK;
    return mela_ctx_default(ctx);
}

struct Guard {
    CtxLock lk;
    explicit Guard(mela_ctx *c) : lk(c->mu) { (void)hipSetDevice(c->device); }
};

// Bytes of the packed transpose (one row of words64(rows) words per column)
// of a rows x cols host array; 0 when the array has no element.
size_t staged_bytes(size_t rows, size_t cols) {
    return rows && cols ? cols * words64(rows) * 8 : 0;
}

// Packs a host column-major array into `packed_slot` (row-major, ld words).
// The host packs its columns into the pinned `stage` (staged_bytes(rows,
// cols) long) in chunks, each chunk's copy to `raw_slot` overlapping the
// packing of the next; one bit-matrix transpose on the device then gives the
// row-major layout.  PCIe carries one bit per element instead of the
// element's 8 (logical) or 64 (double).
mela_status upload_pack(mela_ctx *ctx, const void *host, int cls, size_t rows, size_t cols,
                        uint8_t *stage, int raw_slot, int packed_slot, uint64_t **out,
                        size_t *ld_out) {
    const size_t ld = words64(cols) ? words64(cols) : 1;
    size_t n_el, bytes;
    if (mul_overflows(rows, cols, &n_el) || mul_overflows(n_el, elem_size(cls), &bytes))
        return set_error(MELA_ERR_INVALID, "matrix too large");
    void *packed = nullptr;
    mela_status st = scratch_get(ctx, packed_slot, (rows ? rows : 1) * ld * 8, &packed);
    if (st != MELA_OK) return st;
    if (rows && cols) {
        const size_t ldt = words64(rows);
        void *raw = nullptr;
        if ((st = scratch_get(ctx, raw_slot, staged_bytes(rows, cols), &raw)) != MELA_OK)
            return st;
        // about eight chunks, each at least 8 MiB of the host array
        const size_t col_bytes = rows * elem_size(cls);
        size_t chunk = (cols + 7) / 8;
        const size_t min_cols = (((size_t)8 << 20) + col_bytes - 1) / col_bytes;
        if (chunk < min_cols) chunk = min_cols;
        for (size_t j0 = 0; j0 < cols; j0 += chunk) {
            const size_t nc = cols - j0 < chunk ? cols - j0 : chunk;
            uint64_t *dst = (uint64_t *)stage + j0 * ldt;
            host_pack_columns((const uint8_t *)host + j0 * col_bytes, cls & 0xff, rows, nc, dst);
            MELA_HIP(hipMemcpyAsync((uint64_t *)raw + j0 * ldt, dst, nc * ldt * 8,
                                    hipMemcpyHostToDevice, ctx->stream));
        }
        st = launch_transpose(ctx->stream, (const uint64_t *)raw, ldt, cols, rows,
                              (uint64_t *)packed, ld);
        if (st != MELA_OK) return st;
    } else if (rows) {
        if ((st = launch_fill_zero(ctx->stream, packed, rows * ld * 8)) != MELA_OK) return st;
    }
    *out = (uint64_t *)packed;
    *ld_out = ld;
    return MELA_OK;
}

// The rows x cols packed matrix P (row-major, ldp words) to the host as a
// column-major logical array: the device transposes it into `slot`, the
// packed columns come back through the pinned buffer and the host spreads
// them into bytes.  Returns with the stream idle.
mela_status download_unpack(mela_ctx *ctx, const uint64_t *P, size_t ldp, size_t rows,
                            size_t cols, int slot, uint8_t *host) {
    if (rows && cols) {
        const size_t ldt = words64(rows), tb = staged_bytes(rows, cols);
        void *t = nullptr;
        uint8_t *pin = nullptr;
        mela_status st = scratch_get(ctx, slot, tb, &t);
        if (st != MELA_OK) return st;
        if ((st = launch_transpose(ctx->stream, P, ldp, rows, cols, (uint64_t *)t, ldt)) !=
            MELA_OK)
            return st;
        if ((st = pinned_get(ctx, tb, &pin)) != MELA_OK) return st;
        MELA_HIP(hipMemcpyAsync(pin, t, tb, hipMemcpyDeviceToHost, ctx->stream));
        MELA_HIP(hipStreamSynchronize(ctx->stream));
        host_unpack_columns((const uint64_t *)pin, rows, cols, host);
        return MELA_OK;
    }
    MELA_HIP(hipStreamSynchronize(ctx->stream));
    return MELA_OK;
}

int env_algo(int dflt) {
    const char *e = std::getenv("MELA_MATMUL_ALGO");
    if (!e) return dflt;
    if (!std::strcmp(e, "andpop")) return MELA_ALGO_ANDPOP;
    if (!std::strcmp(e, "m4rm")) return MELA_ALGO_M4RM;
    if (!std::strcmp(e, "mfma_i8")) return MELA_ALGO_MFMA_I8;
    if (!std::strcmp(e, "mfma_fp4")) return MELA_ALGO_MFMA_FP4;
    return dflt;
}

mela_status matmul_entry(mela_ctx *ctx, const void *A, int ac, size_t m, size_t k, const void *B,
                         int bc, size_t k2, size_t n, uint8_t *C, int algo) {
    // matmul_gf2.c:78-88: sparse, inner dimensions, classes -- in that order
    if ((ac & MELA_SPARSE) || (bc & MELA_SPARSE))
        return set_error(MELA_ERR_SPARSE, kSparseMsg);
    if (k != k2) return set_error(MELA_ERR_DIM, "Inner dimensions must agree.");
    if (!is_supported(ac) || !is_supported(bc))
        return set_error(MELA_ERR_TYPE, "Inputs must be logical or double matrices.");
    size_t mn;
    if (mul_overflows(m, n, &mn)) return set_error(MELA_ERR_INVALID, "output too large");
    if ((m * k && !A) || (k * n && !B) || (mn && !C))
        return set_error(MELA_ERR_INVALID, "null data pointer");
    if (mn == 0) return MELA_OK;
    mela_status st0 = resolve(&ctx);
    if (st0 != MELA_OK) return st0;
    Guard g(ctx);
    uint64_t *Ap, *Bp;
    size_t lda, ldb;
    const size_t sa = staged_bytes(m, k);
    uint8_t *pin = nullptr;
    mela_status st = pinned_get(ctx, sa + staged_bytes(k, n), &pin);
    if (st != MELA_OK) return st;
    if ((st = upload_pack(ctx, A, ac, m, k, pin, SLOT_IN_A, SLOT_A, &Ap, &lda)) != MELA_OK)
        return st;
    if ((st = upload_pack(ctx, B, bc, k, n, pin + sa, SLOT_IN_B, SLOT_B, &Bp, &ldb)) != MELA_OK)
        return st;
    const size_t ldc = words64(n);
    void *Cp = nullptr;
    if ((st = scratch_get(ctx, SLOT_C, m * ldc * 8, &Cp)) != MELA_OK) return st;
    st = matmul_device(ctx, Ap, lda, Bp, ldb, (uint64_t *)Cp, ldc, m, k, n, algo);
    if (st != MELA_OK) return st;
    return download_unpack(ctx, (const uint64_t *)Cp, ldc, m, n, SLOT_OUT, C);
}

// ge: the messages of mela_rank_gf2 (else mela_rank_m4ri's); the sweep is
// MELA_ELIM_AUTO for both (the rank does not depend on it)
mela_status rank_entry(mela_ctx *ctx, const void *A, int ac, size_t m, size_t n, double *rank,
                       bool ge) {
    const int variant = MELA_ELIM_AUTO;
    if (!is_supported(ac)) {
        if (ge)  // rank_gf2.c:131-132
            return set_error(MELA_ERR_TYPE, "Input must be a logical or double matrix.",
                             "bitrank_mex:inputNotLogical");
        return set_error(MELA_ERR_TYPE, "Input must be double or logical.");  // rank_m4ri.c:331-334
    }
    // the reference rank gateways do not test for sparse storage; reading it
    // as dense data is not reproducible, so it is rejected like matmul/null
    if (ac & MELA_SPARSE) return set_error(MELA_ERR_SPARSE, kSparseMsg);
    if (!rank) return set_error(MELA_ERR_INVALID, "null argument");
    *rank = 0.0;
    if (m * n && !A) return set_error(MELA_ERR_INVALID, "null data pointer");
    if (m == 0 || n == 0) return MELA_OK;
    mela_status st0 = resolve(&ctx);
    if (st0 != MELA_OK) return st0;
    Guard g(ctx);
    uint64_t *Ap;
    size_t lda;
    uint8_t *pin = nullptr;
    mela_status st = pinned_get(ctx, staged_bytes(m, n), &pin);
    if (st != MELA_OK) return st;
    if ((st = upload_pack(ctx, A, ac, m, n, pin, SLOT_IN_A, SLOT_A, &Ap, &lda)) != MELA_OK)
        return st;
    ElimResult res;
    st = eliminate(ctx, Ap, lda, m, n, variant, false, false, &res, nullptr, nullptr);
    if (st != MELA_OK) return st;
    *rank = (double)res.rank;
    return MELA_OK;
}

// rowmap[p] = pivot index of column p, or -(1 + free index); freecols ascending
void build_null_maps(size_t n, const std::vector<int> &piv, std::vector<int> &rowmap,
                     std::vector<int> &freecols) {
    rowmap.assign(n, 0);
    std::vector<char> isp(n, 0);
    for (size_t i = 0; i < piv.size(); ++i) {
        isp[(size_t)piv[i]] = 1;
        rowmap[(size_t)piv[i]] = (int)i;
    }
    freecols.clear();
    for (size_t c = 0; c < n; ++c)
        if (!isp[c]) {
            rowmap[c] = -1 - (int)freecols.size();
            freecols.push_back((int)c);
        }
}

mela_status upload_maps(mela_ctx *ctx, const std::vector<int> &rowmap,
                        const std::vector<int> &freecols, int **d_rowmap, int **d_free) {
    void *p = nullptr;
    const size_t bytes = (rowmap.size() + freecols.size() + 2) * sizeof(int);
    mela_status st = scratch_get(ctx, SLOT_ELIM_TAB, bytes, &p);
    if (st != MELA_OK) return st;
    int *base = (int *)p;
    if (!rowmap.empty())
        MELA_HIP(hipMemcpyAsync(base, rowmap.data(), rowmap.size() * sizeof(int),
                                hipMemcpyHostToDevice, ctx->stream));
    if (!freecols.empty())
        MELA_HIP(hipMemcpyAsync(base + rowmap.size(), freecols.data(),
                                freecols.size() * sizeof(int), hipMemcpyHostToDevice,
                                ctx->stream));
    *d_rowmap = base;
    *d_free = base + rowmap.size();
    return MELA_OK;
}

mela_status null_entry(mela_ctx *ctx, const void *H, int hc, size_t m, size_t n,
                       mela_alloc_logical_fn alloc, void *user, size_t *nullity) {
    const int variant = MELA_ELIM_AUTO;
    // null_gf2.c:151-156: sparse, then class
    if (hc & MELA_SPARSE) return set_error(MELA_ERR_SPARSE, kSparseMsg);
    if (!is_supported(hc)) return set_error(MELA_ERR_TYPE, "H must be double or logical.");
    if (!alloc) return set_error(MELA_ERR_INVALID, "null argument");
    if (m * n && !H) return set_error(MELA_ERR_INVALID, "null data pointer");
    mela_status st0 = resolve(&ctx);
    if (st0 != MELA_OK) return st0;
    Guard g(ctx);
    std::vector<int> piv;
    ElimResult res;
    uint64_t *Q = nullptr;
    size_t ldq = 1;
    mela_status st;
    if (m && n) {
        uint64_t *Ap;
        size_t lda;
        uint8_t *pin = nullptr;
        if ((st = pinned_get(ctx, staged_bytes(m, n), &pin)) != MELA_OK) return st;
        st = upload_pack(ctx, H, hc, m, n, pin, SLOT_IN_A, SLOT_A, &Ap, &lda);
        if (st != MELA_OK) return st;
        st = eliminate(ctx, Ap, lda, m, n, variant, true, true, &res, &Q, &ldq);
        if (st != MELA_OK) return st;
        piv = res.pivots;
    }
    const size_t f = n - piv.size();
    uint8_t *Z = alloc(user, n, f);
    if (n * f && !Z) return set_error(MELA_ERR_NOMEM, "Out of memory.");
    if (nullity) *nullity = f;
    if (n * f == 0) return MELA_OK;
    std::vector<int> rowmap, freecols;
    build_null_maps(n, piv, rowmap, freecols);
    int *d_rowmap, *d_free;
    if ((st = upload_maps(ctx, rowmap, freecols, &d_rowmap, &d_free)) != MELA_OK) return st;
    // the basis packed (n x f, row-major), then to the host as columns
    const size_t ldz = words64(f);
    void *dz = nullptr;
    if ((st = scratch_get(ctx, SLOT_NULL_Z, n * ldz * 8, &dz)) != MELA_OK) return st;
    if (!Q) {
        void *dummy;
        if ((st = scratch_get(ctx, SLOT_ELIM_Q, 8, &dummy)) != MELA_OK) return st;
        Q = (uint64_t *)dummy;
    }
    if ((st = launch_null_packed(ctx, Q, ldq, d_rowmap, piv, n, f, (uint64_t *)dz, ldz)) !=
        MELA_OK)
        return st;
    return download_unpack(ctx, (const uint64_t *)dz, ldz, n, f, SLOT_OUT, Z);
}

mela_status check_variant(int variant) {
    if (variant < MELA_ELIM_GE || variant > MELA_ELIM_PANEL)
        return set_error(MELA_ERR_INVALID, "unknown elimination variant");
    return MELA_OK;
}

mela_status check_bitmat(const mela_bitmat *M, const char *name) {
    if (!M) return set_error(MELA_ERR_INVALID, std::string("null descriptor ") + name);
    if (M->rows && M->cols && !M->data)
        return set_error(MELA_ERR_INVALID, std::string("null data in ") + name);
    if (M->cols && M->ld < words64(M->cols))
        return set_error(MELA_ERR_INVALID, std::string("ld too small in ") + name);
    return MELA_OK;
}

}  // namespace

extern "C" {

mela_status mela_matmul_gf2(mela_ctx *ctx, const void *A, int a_class, size_t m, size_t k,
                            const void *B, int b_class, size_t k2, size_t n, uint8_t *C) {
    return matmul_entry(ctx, A, a_class, m, k, B, b_class, k2, n, C, env_algo(MELA_ALGO_AUTO));
}

// Both product gateways run AUTO: the product does not depend on the
// algorithm (the reference's M4RI gateway differs from its GE one only in
// speed), so the M4RI one takes the fastest engine too; k_m4rm stays
// available through mela_matmul_packed (MELA_ALGO_M4RM) or
// MELA_MATMUL_ALGO=m4rm.
mela_status mela_matmul_m4ri(mela_ctx *ctx, const void *A, int a_class, size_t m, size_t k,
                             const void *B, int b_class, size_t k2, size_t n, uint8_t *C) {
    return matmul_entry(ctx, A, a_class, m, k, B, b_class, k2, n, C, env_algo(MELA_ALGO_AUTO));
}

mela_status mela_rank_gf2(mela_ctx *ctx, const void *A, int a_class, size_t m, size_t n,
                          double *rank) {
    return rank_entry(ctx, A, a_class, m, n, rank, true);
}

mela_status mela_rank_m4ri(mela_ctx *ctx, const void *A, int a_class, size_t m, size_t n,
                           double *rank) {
    return rank_entry(ctx, A, a_class, m, n, rank, false);
}

mela_status mela_null_gf2(mela_ctx *ctx, const void *H, int h_class, size_t m, size_t n,
                          mela_alloc_logical_fn alloc, void *user, size_t *nullity) {
    return null_entry(ctx, H, h_class, m, n, alloc, user, nullity);
}

mela_status mela_null_m4ri(mela_ctx *ctx, const void *H, int h_class, size_t m, size_t n,
                           mela_alloc_logical_fn alloc, void *user, size_t *nullity) {
    return null_entry(ctx, H, h_class, m, n, alloc, user, nullity);
}

mela_status mela_rref_gf2(mela_ctx *ctx, const void *A, int a_class, size_t m, size_t n,
                          uint8_t *R, double *jb, size_t *rank) {
    // gf2rref.m:49-73 semantics; messages follow the rank entry points
    if (a_class & MELA_SPARSE) return set_error(MELA_ERR_SPARSE, kSparseMsg);
    if (!is_supported(a_class))
        return set_error(MELA_ERR_TYPE, "Input must be a logical or double matrix.");
    if (!rank) return set_error(MELA_ERR_INVALID, "null argument");
    *rank = 0;
    size_t mn;
    if (mul_overflows(m, n, &mn)) return set_error(MELA_ERR_INVALID, "matrix too large");
    if (mn && (!A || !R)) return set_error(MELA_ERR_INVALID, "null data pointer");
    if (mn == 0) return MELA_OK;
    mela_status st = resolve(&ctx);
    if (st != MELA_OK) return st;
    Guard g(ctx);
    uint64_t *Ap;
    size_t lda;
    uint8_t *pin = nullptr;
    if ((st = pinned_get(ctx, staged_bytes(m, n), &pin)) != MELA_OK) return st;
    if ((st = upload_pack(ctx, A, a_class, m, n, pin, SLOT_IN_A, SLOT_A, &Ap, &lda)) != MELA_OK)
        return st;
    ElimResult res;
    uint64_t *Q = nullptr;
    size_t ldq = 1;
    if ((st = eliminate(ctx, Ap, lda, m, n, MELA_ELIM_AUTO, true, true, &res, &Q, &ldq)) != MELA_OK)
        return st;
    // R: the rank pivot rows on top, zero rows below, unpacked column-major
    const size_t wc = words64(n);
    void *Rp = nullptr;
    if ((st = scratch_get(ctx, SLOT_C, m * wc * 8, &Rp)) != MELA_OK) return st;
    if ((st = launch_fill_zero(ctx->stream, Rp, m * wc * 8)) != MELA_OK) return st;
    if (res.rank)
        MELA_HIP(hipMemcpy2DAsync(Rp, wc * 8, Q, ldq * 8, wc * 8, res.rank, hipMemcpyDeviceToDevice,
                                  ctx->stream));
    if ((st = download_unpack(ctx, (const uint64_t *)Rp, wc, m, n, SLOT_OUT, R)) != MELA_OK)
        return st;
    if (jb)
        for (size_t i = 0; i < res.rank; ++i) jb[i] = (double)res.pivots[i] + 1.0;
    *rank = res.rank;
    return MELA_OK;
}

mela_status mela_pack_csc(mela_ctx *ctx, const uint64_t *jc, const uint64_t *ir,
                          const void *values, int values_class, size_t m, size_t n,
                          mela_bitmat *A) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    mela_status st = check_bitmat(A, "A");
    if (st != MELA_OK) return st;
    if (A->rows != m || A->cols != n) return set_error(MELA_ERR_DIM, "A must be m x n");
    if (!jc) return set_error(MELA_ERR_INVALID, "null column pointer array");
    if (values && !is_supported(values_class))
        return set_error(MELA_ERR_TYPE, "Inputs must be logical or double matrices.");
    // the column pointers must describe nnz entries in order (jc[0] = 0,
    // non-decreasing); every entry then belongs to exactly one column
    if (jc[0] != 0) return set_error(MELA_ERR_INVALID, "jc[0] must be 0");
    for (size_t j = 0; j < n; ++j)
        if (jc[j + 1] < jc[j]) return set_error(MELA_ERR_INVALID, "jc must be non-decreasing");
    const size_t nnz = jc[n];
    if (nnz && !ir) return set_error(MELA_ERR_INVALID, "null row index array");
    Guard g(ctx);
    if (m && A->ld && (st = launch_fill_zero(ctx->stream, A->data, m * A->ld * 8)) != MELA_OK)
        return st;
    if (nnz == 0) {
        MELA_HIP(hipStreamSynchronize(ctx->stream));
        return MELA_OK;
    }
    const int vcls = values ? (values_class & 0xff) : -1;
    const size_t vbytes = values ? nnz * elem_size(values_class) : 0;
    const size_t off_ir = (n + 1) * 8, off_v = off_ir + nnz * 8;
    const size_t off_err = (off_v + vbytes + 7) / 8 * 8;
    void *buf = nullptr;
    if ((st = scratch_get(ctx, SLOT_IN_A, off_err + 8, &buf)) != MELA_OK) return st;
    char *b = (char *)buf;
    MELA_HIP(hipMemcpyAsync(b, jc, (n + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    MELA_HIP(hipMemcpyAsync(b + off_ir, ir, nnz * 8, hipMemcpyHostToDevice, ctx->stream));
    if (vbytes) MELA_HIP(hipMemcpyAsync(b + off_v, values, vbytes, hipMemcpyHostToDevice, ctx->stream));
    if ((st = launch_fill_zero(ctx->stream, b + off_err, 8)) != MELA_OK) return st;
    if ((st = launch_pack_csc(ctx->stream, (const uint64_t *)b, (const uint64_t *)(b + off_ir),
                              vbytes ? b + off_v : nullptr, vcls, nnz, m, n, A->data, A->ld,
                              (unsigned *)(b + off_err))) != MELA_OK)
        return st;
    unsigned err = 0;
    MELA_HIP(hipMemcpyAsync(&err, b + off_err, sizeof(err), hipMemcpyDeviceToHost, ctx->stream));
    MELA_HIP(hipStreamSynchronize(ctx->stream));
    if (err) return set_error(MELA_ERR_INVALID, "row index out of range in ir");
    return MELA_OK;
}

mela_status mela_pack_dev(mela_ctx *ctx, const void *src_dev, int src_class, mela_bitmat *M) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    mela_status st = check_bitmat(M, "M");
    if (st != MELA_OK) return st;
    if (!is_supported(src_class)) return set_error(MELA_ERR_TYPE, "Inputs must be logical or double matrices.");
    Guard g(ctx);
    return launch_pack(ctx->stream, src_dev, src_class & 0xff, M->rows, M->cols, M->data, M->ld);
}

mela_status mela_unpack_dev(mela_ctx *ctx, const mela_bitmat *M, uint8_t *dst_dev) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    mela_status st = check_bitmat(M, "M");
    if (st != MELA_OK) return st;
    Guard g(ctx);
    return launch_unpack(ctx->stream, M->data, M->ld, M->rows, M->cols, dst_dev);
}

mela_status mela_matmul_packed(mela_ctx *ctx, const mela_bitmat *A, const mela_bitmat *B,
                               mela_bitmat *C, int algo) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    mela_status st;
    if ((st = check_bitmat(A, "A")) || (st = check_bitmat(B, "B")) || (st = check_bitmat(C, "C")))
        return st;
    if (A->cols != B->rows) return set_error(MELA_ERR_DIM, "Inner dimensions must agree.");
    if (C->rows != A->rows || C->cols != B->cols)
        return set_error(MELA_ERR_DIM, "output shape must be A.rows x B.cols");
    Guard g(ctx);
    return matmul_device(ctx, A->data, A->ld, B->data, B->ld, C->data, C->ld, A->rows, A->cols,
                         B->cols, algo);
}

mela_status mela_rank_packed(mela_ctx *ctx, const mela_bitmat *A, int variant, size_t *rank) {
    if (!ctx || !rank) return set_error(MELA_ERR_INVALID, "null argument");
    mela_status st = check_bitmat(A, "A");
    if (st != MELA_OK || (st = check_variant(variant)) != MELA_OK) return st;
    Guard g(ctx);
    ElimResult res;
    st = eliminate(ctx, A->data, A->ld, A->rows, A->cols, variant, false, false, &res, nullptr,
                   nullptr);
    if (st != MELA_OK) return st;
    *rank = res.rank;
    return MELA_OK;
}

mela_status mela_rref_packed(mela_ctx *ctx, const mela_bitmat *A, int variant, mela_bitmat *R,
                             int *pivots, size_t *rank) {
    if (!ctx) return set_error(MELA_ERR_INVALID, "null context");
    mela_status st = check_bitmat(A, "A");
    if (st != MELA_OK || (st = check_variant(variant)) != MELA_OK) return st;
    const size_t rmax = A->rows < A->cols ? A->rows : A->cols;
    if (R) {
        if ((st = check_bitmat(R, "R")) != MELA_OK) return st;
        if (R->rows < rmax || R->cols != A->cols)
            return set_error(MELA_ERR_DIM, "R must be min(rows,cols) x cols");
    }
    Guard g(ctx);
    ElimResult res;
    uint64_t *Q = nullptr;
    size_t ldq = 1;
    st = eliminate(ctx, A->data, A->ld, A->rows, A->cols, variant, true, true, &res, &Q, &ldq);
    if (st != MELA_OK) return st;
    if (rank) *rank = res.rank;
    if (pivots && res.rank) std::memcpy(pivots, res.pivots.data(), res.rank * sizeof(int));
    if (R && R->rows) {
        if ((st = launch_fill_zero(ctx->stream, R->data, R->rows * R->ld * 8)) != MELA_OK) return st;
        if (res.rank)
            MELA_HIP(hipMemcpy2DAsync(R->data, R->ld * 8, Q, ldq * 8, words64(A->cols) * 8, res.rank,
                                      hipMemcpyDeviceToDevice, ctx->stream));
        MELA_HIP(hipStreamSynchronize(ctx->stream));
    }
    return MELA_OK;
}

mela_status mela_null_packed(mela_ctx *ctx, const mela_bitmat *A, int variant, mela_bitmat *Z,
                             size_t *nullity) {
    if (!ctx || !Z) return set_error(MELA_ERR_INVALID, "null argument");
    mela_status st = check_bitmat(A, "A");
    if (st != MELA_OK || (st = check_variant(variant)) != MELA_OK) return st;
    const size_t n = A->cols;
    std::vector<int> piv;
    uint64_t *Q = nullptr;
    size_t ldq = 1;
    // one hold of the context over elimination and extraction: the pivot
    // rows Q and the maps are context scratch that another call on this
    // context would overwrite (mela_bitmat_alloc takes no context lock)
    Guard g(ctx);
    if (A->rows && n) {
        ElimResult res;
        st = eliminate(ctx, A->data, A->ld, A->rows, n, variant, true, true, &res, &Q, &ldq);
        if (st != MELA_OK) return st;
        piv = res.pivots;
    }
    const size_t f = n - piv.size();
    if ((st = mela_bitmat_alloc(ctx, n, f, Z)) != MELA_OK) return st;
    if (nullity) *nullity = f;
    if (n == 0 || f == 0) return MELA_OK;
    std::vector<int> rowmap, freecols;
    build_null_maps(n, piv, rowmap, freecols);
    int *d_rowmap, *d_free;
    if ((st = upload_maps(ctx, rowmap, freecols, &d_rowmap, &d_free)) != MELA_OK) return st;
    if (!Q) {
        void *dummy;
        if ((st = scratch_get(ctx, SLOT_ELIM_Q, 8, &dummy)) != MELA_OK) return st;
        Q = (uint64_t *)dummy;
    }
    st = launch_null_packed(ctx, Q, ldq, d_rowmap, piv, n, f, Z->data, Z->ld);
    if (st != MELA_OK) return st;
    return MELA_OK;
}

}  // extern "C"
