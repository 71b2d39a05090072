/* mex_shim.c -- implementation of the mex.h stand-in (see mex.h). */
#include "mex.h"

#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct mxArray_tag {
    int cls;
    int sparse;
    mwSize m, n;
    void *data;  /* SHIM_COMPLEX: real parts, or (re, im) pairs when interleaved */
    double *pi;  /* SHIM_COMPLEX with separate storage: imaginary parts */
};

static __thread jmp_buf *t_jmp;
static __thread char t_msg[1024];
static __thread char t_id[256];

static size_t esize(int cls) {
    if (cls == SHIM_COMPLEX) return MX_HAS_INTERLEAVED_COMPLEX ? 16 : 8;
    return cls == SHIM_DOUBLE ? 8 : (cls == SHIM_INT32 ? 4 : 1);
}

mwSize mxGetM(const mxArray *a) { return a->m; }
mwSize mxGetN(const mxArray *a) { return a->n; }
bool mxIsSparse(const mxArray *a) { return a->sparse != 0; }
bool mxIsLogical(const mxArray *a) { return a->cls == SHIM_LOGICAL; }
bool mxIsDouble(const mxArray *a) { return a->cls == SHIM_DOUBLE || a->cls == SHIM_COMPLEX; }
bool mxIsComplex(const mxArray *a) { return a->cls == SHIM_COMPLEX; }
void *mxGetData(const mxArray *a) { return a->data; }
double *mxGetPr(const mxArray *a) { return mxIsDouble(a) ? (double *)a->data : NULL; }
mxLogical *mxGetLogicals(const mxArray *a) {
    return a->cls == SHIM_LOGICAL ? (mxLogical *)a->data : NULL;
}
void *mxCalloc(size_t n, size_t size) { return calloc(n ? n : 1, size ? size : 1); }
void mxFree(void *p) { free(p); }

mxArray *shim_create(int cls, mwSize m, mwSize n, int sparse, const void *data) {
    mxArray *a = (mxArray *)calloc(1, sizeof *a);
    a->cls = cls;
    a->sparse = sparse;
    a->m = m;
    a->n = n;
    a->data = calloc((m * n) != 0 ? m * n : 1, esize(cls));
    a->pi = NULL;
    if (cls == SHIM_COMPLEX && !MX_HAS_INTERLEAVED_COMPLEX) {
        /* separate storage: split the (re, im) pairs */
        a->pi = (double *)calloc((m * n) != 0 ? m * n : 1, sizeof(double));
        const double *src = (const double *)data;
        for (size_t i = 0; data && i < m * n; ++i) {
            ((double *)a->data)[i] = src[2 * i];
            a->pi[i] = src[2 * i + 1];
        }
    } else if (data && (m * n) != 0) {
        memcpy(a->data, data, m * n * esize(cls));
    }
    return a;
}

void shim_destroy(mxArray *a) {
    if (!a) return;
    free(a->data);
    free(a->pi);
    free(a);
}

mxArray *mxCreateLogicalMatrix(mwSize m, mwSize n) { return shim_create(SHIM_LOGICAL, m, n, 0, NULL); }

mxArray *mxCreateDoubleScalar(double v) { return shim_create(SHIM_DOUBLE, 1, 1, 0, &v); }

static void raise_(void) {
    if (t_jmp) longjmp(*t_jmp, 1);
    fprintf(stderr, "mex error outside shim_call: %s\n", t_msg);
    abort();
}

void mexErrMsgTxt(const char *msg) {
    snprintf(t_msg, sizeof t_msg, "%s", msg);
    t_id[0] = 0;
    raise_();
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_msg, sizeof t_msg, fmt, ap);
    va_end(ap);
    snprintf(t_id, sizeof t_id, "%s", id);
    raise_();
}

int shim_call(shim_mexfn fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs) {
    jmp_buf jb;
    jmp_buf *prev = t_jmp;
    t_msg[0] = 0;
    t_id[0] = 0;
    t_jmp = &jb;
    if (setjmp(jb)) {
        t_jmp = prev;
        return 1;
    }
    fn(nlhs, plhs, nrhs, prhs);
    t_jmp = prev;
    return 0;
}

const char *shim_last_msg(void) { return t_msg; }
const char *shim_last_id(void) { return t_id; }
int shim_class(const mxArray *a) { return a->cls; }
