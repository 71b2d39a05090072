/*
 * mela_gateway.c -- MEX gateways of the six functions, over the C ABI.
 *
 * Compiled once per function with -DMELA_FN_<NAME> into its own MEX file
 * (mela_matmul_gf2, mela_matmul_m4ri, mela_rank_gf2, mela_rank_m4ri,
 * mela_null_gf2, mela_null_m4ri), replacing mexFunction at
 * src/mela_matmul_gf2.c:72, src/mela_matmul_m4ri.c:140, src/mela_rank_gf2.c:126,
 * src/mela_rank_m4ri.c:286, src/mela_null_gf2.c:142 and src/mela_null_m4ri.c:251.
 * Arity checks and their messages are the reference's; argument validation and
 * the work happen in libmela (which keeps the reference's validation order and
 * text), and a failure is re-raised with mexErrMsgIdAndTxt/mexErrMsgTxt.
 * Complex double inputs count by their real parts under either MATLAB array
 * API (real_data).
 */
#include "mex.h"
#include "mela.h"

static int mx_class(const mxArray *a) {
    int c = mxIsLogical(a) ? MELA_LOGICAL : (mxIsDouble(a) ? MELA_DOUBLE : MELA_OTHER);
    if (mxIsSparse(a)) c |= MELA_SPARSE;
    return c;
}

/* The elements the library reads.  A complex double array counts by its real
 * parts, as in the reference, whose default (separate-complex) build reads
 * them through mxGetData / mxGetPr (e.g. null_gf2.c:170, rank_m4ri.c:309).
 * Under the interleaved complex API (MX_HAS_INTERLEAVED_COMPLEX, `mex
 * -R2018a`) mxGetData returns (re, im) pairs, so the real parts are gathered
 * into *tmp first (released by the caller with mxFree). */
static const void *real_data(const mxArray *a, void **tmp) {
    *tmp = NULL;
#if MX_HAS_INTERLEAVED_COMPLEX
    if (mxIsDouble(a) && mxIsComplex(a) && !mxIsSparse(a)) {
        const size_t cnt = (size_t)mxGetM(a) * (size_t)mxGetN(a);
        const double *src = (const double *)mxGetData(a);
        double *re = (double *)mxCalloc(cnt ? cnt : 1, sizeof(double));
        if (!re) mexErrMsgTxt("Out of memory.");
        for (size_t i = 0; i < cnt; ++i) re[i] = src[2 * i];
        *tmp = re;
        return re;
    }
#endif
    return mxGetData(a);
}

static void raise_status(void) {
    const char *id = mela_last_error_id();
    if (id && id[0])
        mexErrMsgIdAndTxt(id, "%s", mela_last_error());
    else
        mexErrMsgTxt(mela_last_error());
}

#if defined(MELA_FN_MATMUL_GF2) || defined(MELA_FN_MATMUL_M4RI)
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs != 2) mexErrMsgTxt("Need two input matrices A,B.");
    if (nlhs > 1) mexErrMsgTxt("Too many outputs.");
    mwSize m = mxGetM(prhs[0]), k = mxGetN(prhs[0]);
    mwSize k2 = mxGetM(prhs[1]), n = mxGetN(prhs[1]);
    int ca = mx_class(prhs[0]), cb = mx_class(prhs[1]);
    /* validation first (no output is created for a rejected call) */
    mela_status st = MELA_OK;
    if ((ca & MELA_SPARSE) || (cb & MELA_SPARSE) || k != k2 ||
        (ca & 0xff) == MELA_OTHER || (cb & 0xff) == MELA_OTHER) {
#ifdef MELA_FN_MATMUL_GF2
        st = mela_matmul_gf2(NULL, NULL, ca, m, k, NULL, cb, k2, n, NULL);
#else
        st = mela_matmul_m4ri(NULL, NULL, ca, m, k, NULL, cb, k2, n, NULL);
#endif
        raise_status();
    }
    plhs[0] = mxCreateLogicalMatrix(m, n);
    void *ta, *tb;
    const void *da = real_data(prhs[0], &ta), *db = real_data(prhs[1], &tb);
#ifdef MELA_FN_MATMUL_GF2
    st = mela_matmul_gf2(NULL, da, ca, m, k, db, cb, k2, n, (uint8_t *)mxGetLogicals(plhs[0]));
#else
    st = mela_matmul_m4ri(NULL, da, ca, m, k, db, cb, k2, n, (uint8_t *)mxGetLogicals(plhs[0]));
#endif
    if (ta) mxFree(ta);
    if (tb) mxFree(tb);
    if (st != MELA_OK) raise_status();
}
#endif

#if defined(MELA_FN_RANK_GF2) || defined(MELA_FN_RANK_M4RI)
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    (void)nlhs;
#ifdef MELA_FN_RANK_GF2
    if (nrhs != 1) mexErrMsgIdAndTxt("bitrank_mex:nargin", "One input required.");
#else
    if (nrhs != 1) mexErrMsgTxt("One input required.");
#endif
    double r = 0.0;
    mela_status st;
    void *t;
    const void *d = real_data(prhs[0], &t);
#ifdef MELA_FN_RANK_GF2
    st = mela_rank_gf2(NULL, d, mx_class(prhs[0]), mxGetM(prhs[0]), mxGetN(prhs[0]), &r);
#else
    st = mela_rank_m4ri(NULL, d, mx_class(prhs[0]), mxGetM(prhs[0]), mxGetN(prhs[0]), &r);
#endif
    if (t) mxFree(t);
    if (st != MELA_OK) raise_status();
    plhs[0] = mxCreateDoubleScalar(r);
}
#endif

#if defined(MELA_FN_NULL_GF2) || defined(MELA_FN_NULL_M4RI)
/* creates the n x nullity output once its size is known (null_gf2.c:208) */
static uint8_t *alloc_out(void *user, size_t rows, size_t cols) {
    mxArray **slot = (mxArray **)user;
    *slot = mxCreateLogicalMatrix(rows, cols);
    return (uint8_t *)mxGetLogicals(*slot);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs != 1) mexErrMsgTxt("One input required: H.");
    if (nlhs > 1) mexErrMsgTxt("One output required.");
    size_t f = 0;
    mela_status st;
    mxArray *out = NULL;
    void *t;
    const void *d = real_data(prhs[0], &t);
#ifdef MELA_FN_NULL_GF2
    st = mela_null_gf2(NULL, d, mx_class(prhs[0]), mxGetM(prhs[0]), mxGetN(prhs[0]), alloc_out,
                       &out, &f);
#else
    st = mela_null_m4ri(NULL, d, mx_class(prhs[0]), mxGetM(prhs[0]), mxGetN(prhs[0]), alloc_out,
                        &out, &f);
#endif
    if (t) mxFree(t);
    if (st != MELA_OK) raise_status();
    plhs[0] = out;
}
#endif
