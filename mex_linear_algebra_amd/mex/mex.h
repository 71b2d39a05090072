/*
 * mex.h -- minimal stand-in for MATLAB's MEX API, for building and testing the
 * gateways in this directory without MATLAB.  It declares exactly the 17
 * names the reference gateways use (mwSize, mxLogical, mxArray, mxGetM,
 * mxGetN, mxIsSparse, mxIsLogical, mxIsDouble, mxGetData, mxGetPr,
 * mxGetLogicals, mxCalloc, mxFree, mxCreateLogicalMatrix,
 * mxCreateDoubleScalar, mexErrMsgTxt, mexErrMsgIdAndTxt), mxIsComplex for
 * complex doubles, plus test helpers (shim_*) to create inputs and call a
 * gateway with error capture.  With a real MATLAB installation the gateways
 * compile against MATLAB's own mex.h.
 *
 * Complex doubles follow MATLAB's two array APIs, chosen like MATLAB's mex.h
 * by MX_HAS_INTERLEAVED_COMPLEX: 0 (the default `mex` build, separate
 * complex storage): mxGetData / mxGetPr return the real parts; 1 (`mex
 * -R2018a`, interleaved complex): mxGetData returns (re, im) pairs.
 */
#ifndef MELA_MEX_SHIM_H
#define MELA_MEX_SHIM_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef MX_HAS_INTERLEAVED_COMPLEX
#define MX_HAS_INTERLEAVED_COMPLEX 0
#endif

typedef size_t mwSize;
typedef bool mxLogical;
typedef struct mxArray_tag mxArray;

mwSize mxGetM(const mxArray *a);
mwSize mxGetN(const mxArray *a);
bool mxIsSparse(const mxArray *a);
bool mxIsLogical(const mxArray *a);
bool mxIsDouble(const mxArray *a);
bool mxIsComplex(const mxArray *a);
void *mxGetData(const mxArray *a);
double *mxGetPr(const mxArray *a);
mxLogical *mxGetLogicals(const mxArray *a);
void *mxCalloc(size_t n, size_t size);
void mxFree(void *p);
mxArray *mxCreateLogicalMatrix(mwSize m, mwSize n);
mxArray *mxCreateDoubleScalar(double v);
void mexErrMsgTxt(const char *msg);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);

/* ---- test helpers (not part of MATLAB's API) ---- */
enum { SHIM_LOGICAL = 0, SHIM_DOUBLE = 1, SHIM_INT32 = 2, SHIM_COMPLEX = 3 };
/* data: column-major elements; for SHIM_COMPLEX m*n (re, im) pairs */
mxArray *shim_create(int cls, mwSize m, mwSize n, int sparse, const void *data);
void shim_destroy(mxArray *a);
typedef void (*shim_mexfn)(int, mxArray **, int, const mxArray **);
/* runs fn; returns 0, or 1 if it raised (message/id then readable below) */
int shim_call(shim_mexfn fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs);
const char *shim_last_msg(void);
const char *shim_last_id(void);
int shim_class(const mxArray *a);

#ifdef __cplusplus
}
#endif
#endif
