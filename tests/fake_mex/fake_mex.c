/*
 * fake_mex.c -- test-only host for the MEX gateway (no MATLAB in the image).
 *
 * Emulates what the gateway relies on:
 *   - mxArray: full double arrays (real or interleaved complex, up to 8
 *     dimensions, column-major), real sparse double matrices (CSC with
 *     mwIndex jc/ir) and char rows;
 *   - mexErrMsgIdAndTxt does not return: it longjmps back to fm_call, which
 *     reports the identifier and message and discards the outputs;
 *   - mxMalloc memory still held when the gateway returns is released, as
 *     MATLAB does at the end of every MEX call; fm_leaked() counts the blocks
 *     a call that succeeded left to that release (a tidy gateway leaves none);
 *   - mexLock / mexUnlock / mexAtExit bookkeeping; fm_exit() runs the exit
 *     handler as MATLAB does when the MEX file is cleared or MATLAB quits.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

enum { CLS_DOUBLE = 0, CLS_CHAR = 1 };

struct mxArray_tag {
  int cls, cplx, sparse;
  mwSize ndim;
  mwSize dims[8];
  double* data; /* full: numel (x2 when complex); sparse: pr[nzmax] */
  mwIndex *jc, *ir;
  char* str;
};

static jmp_buf jb;
static int in_call;
static char err_id[128], err_msg[1024];
static void** blocks;
static size_t nblocks, capblocks;
static long long leaked;
static int lock_count;
static void (*exit_fn)(void);

/* ---- mx* ---------------------------------------------------------------- */

static size_t numel(const mxArray* a) {
  size_t n = 1;
  for (mwSize i = 0; i < a->ndim; ++i) n *= a->dims[i];
  return n;
}

size_t mxGetM(const mxArray* a) { return a->dims[0]; }
size_t mxGetN(const mxArray* a) {
  size_t n = 1;
  for (mwSize i = 1; i < a->ndim; ++i) n *= a->dims[i];
  return n;
}
size_t mxGetNumberOfElements(const mxArray* a) { return numel(a); }
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->ndim; }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims; }
bool mxIsDouble(const mxArray* a) { return a->cls == CLS_DOUBLE; }
bool mxIsComplex(const mxArray* a) { return a->cls == CLS_DOUBLE && a->cplx; }
bool mxIsSparse(const mxArray* a) { return a->sparse != 0; }
bool mxIsChar(const mxArray* a) { return a->cls == CLS_CHAR; }
bool mxIsEmpty(const mxArray* a) { return numel(a) == 0; }

int mxGetString(const mxArray* a, char* buf, mwSize buflen) {
  if (a->cls != CLS_CHAR || buflen == 0) return 1;
  const size_t len = strlen(a->str);
  const size_t k = len < buflen - 1 ? len : buflen - 1;
  memcpy(buf, a->str, k);
  buf[k] = 0;
  return len < buflen ? 0 : 1;
}

double mxGetScalar(const mxArray* a) {
  if (a->cls != CLS_DOUBLE || !a->data) return 0.0;
  if (a->sparse) return a->jc[a->dims[1]] > 0 ? a->data[0] : 0.0;
  return numel(a) ? a->data[0] : 0.0;
}

mxDouble* mxGetDoubles(const mxArray* a) { return (a->cls == CLS_DOUBLE && !a->cplx) ? a->data : NULL; }
mxComplexDouble* mxGetComplexDoubles(const mxArray* a) {
  return (a->cls == CLS_DOUBLE && a->cplx) ? (mxComplexDouble*)a->data : NULL;
}
mwIndex* mxGetIr(const mxArray* a) { return a->sparse ? a->ir : NULL; }
mwIndex* mxGetJc(const mxArray* a) { return a->sparse ? a->jc : NULL; }

static mxArray* new_array(void) {
  mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
  if (!a) abort();
  return a;
}

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
  mxArray* a = new_array();
  a->cls = CLS_DOUBLE;
  a->cplx = c == mxCOMPLEX;
  a->ndim = 2;
  a->dims[0] = m;
  a->dims[1] = n;
  a->data = (double*)calloc(m * n * (a->cplx ? 2 : 1) + 1, sizeof(double));
  if (!a->data) abort();
  return a;
}

mxArray* mxCreateDoubleScalar(double v) {
  mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
  a->data[0] = v;
  return a;
}

void mxDestroyArray(mxArray* a) {
  if (!a) return;
  free(a->data);
  free(a->jc);
  free(a->ir);
  free(a->str);
  free(a);
}

void* mxMalloc(mwSize n) {
  void* p = malloc(n ? n : 1);
  if (!p) mexErrMsgIdAndTxt("MATLAB:nomem", "out of memory");
  if (nblocks == capblocks) {
    capblocks = capblocks ? 2 * capblocks : 64;
    blocks = (void**)realloc(blocks, capblocks * sizeof(void*));
    if (!blocks) abort();
  }
  blocks[nblocks++] = p;
  return p;
}

void* mxCalloc(mwSize n, mwSize size) {
  void* p = mxMalloc(n * size);
  memset(p, 0, n * size);
  return p;
}

void mxFree(void* p) {
  if (!p) return;
  for (size_t i = 0; i < nblocks; ++i)
    if (blocks[i] == p) {
      blocks[i] = blocks[--nblocks];
      free(p);
      return;
    }
  fprintf(stderr, "fake_mex: mxFree of a pointer mxMalloc did not return\n");
  abort();
}

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  snprintf(err_id, sizeof err_id, "%s", id ? id : "");
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(err_msg, sizeof err_msg, fmt, ap);
  va_end(ap);
  if (!in_call) {
    fprintf(stderr, "fake_mex: error outside a MEX call: %s: %s\n", err_id, err_msg);
    abort();
  }
  longjmp(jb, 1);
}

int mexPrintf(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  const int r = vprintf(fmt, ap);
  va_end(ap);
  return r;
}

int mexAtExit(void (*fn)(void)) {
  exit_fn = fn;
  return 0;
}
void mexLock(void) { ++lock_count; }
void mexUnlock(void) {
  if (lock_count > 0) --lock_count;
}
bool mexIsLocked(void) { return lock_count > 0; }

/* ---- harness (called from the tests through ctypes) ------------------------ */

mxArray* fm_dense(int ndim, const int64_t* dims, int cplx, const double* data) {
  if (ndim < 2 || ndim > 8) return NULL;
  mxArray* a = new_array();
  a->cls = CLS_DOUBLE;
  a->cplx = cplx != 0;
  a->ndim = (mwSize)ndim;
  for (int i = 0; i < ndim; ++i) a->dims[i] = (mwSize)dims[i];
  const size_t cnt = numel(a) * (a->cplx ? 2 : 1);
  a->data = (double*)calloc(cnt + 1, sizeof(double));
  if (!a->data) abort();
  if (data && cnt) memcpy(a->data, data, cnt * sizeof(double));
  return a;
}

mxArray* fm_sparse(int64_t m, int64_t n, const int64_t* jc, const int64_t* ir, const double* pr) {
  mxArray* a = new_array();
  a->cls = CLS_DOUBLE;
  a->sparse = 1;
  a->ndim = 2;
  a->dims[0] = (mwSize)m;
  a->dims[1] = (mwSize)n;
  const int64_t nnz = jc[n];
  a->jc = (mwIndex*)malloc((size_t)(n + 1) * sizeof(mwIndex));
  a->ir = (mwIndex*)malloc((size_t)(nnz + 1) * sizeof(mwIndex));
  a->data = (double*)malloc((size_t)(nnz + 1) * sizeof(double));
  if (!a->jc || !a->ir || !a->data) abort();
  for (int64_t j = 0; j <= n; ++j) a->jc[j] = (mwIndex)jc[j];
  for (int64_t k = 0; k < nnz; ++k) {
    a->ir[k] = (mwIndex)ir[k];
    a->data[k] = pr[k];
  }
  return a;
}

mxArray* fm_string(const char* s) {
  mxArray* a = new_array();
  a->cls = CLS_CHAR;
  a->ndim = 2;
  a->dims[0] = 1;
  a->dims[1] = strlen(s);
  const size_t len = strlen(s);
  a->str = (char*)malloc(len + 1);
  if (!a->str) abort();
  memcpy(a->str, s, len + 1);
  return a;
}

void fm_destroy(mxArray* a) { mxDestroyArray(a); }

/* 0: the gateway returned; 1: it raised an error (fm_error_id / fm_error_msg),
 * its outputs were discarded.  plhs must hold max(nlhs, 1) entries. */
int fm_call(int nlhs, mxArray** plhs, int nrhs, mxArray** prhs) {
  const int nout = nlhs > 0 ? nlhs : 1;
  for (int i = 0; i < nout; ++i) plhs[i] = NULL;
  err_id[0] = err_msg[0] = 0;
  in_call = 1;
  volatile int failed = 0;
  if (setjmp(jb) == 0) {
    mexFunction(nlhs, plhs, nrhs, (const mxArray**)prhs);
  } else {
    failed = 1;
  }
  in_call = 0;
  if (!failed) leaked += (long long)nblocks;
  while (nblocks) free(blocks[--nblocks]);
  if (failed)
    for (int i = 0; i < nout; ++i) {
      mxDestroyArray(plhs[i]);
      plhs[i] = NULL;
    }
  return failed;
}

const char* fm_error_id(void) { return err_id; }
const char* fm_error_msg(void) { return err_msg; }

/* shape of an output: dims (up to 8), returns ndim; *cplx set */
int fm_shape(const mxArray* a, int64_t* dims, int* cplx) {
  for (mwSize i = 0; i < a->ndim; ++i) dims[i] = (int64_t)a->dims[i];
  *cplx = a->cplx;
  return (int)a->ndim;
}

/* data of a full double array: numel doubles, or 2*numel interleaved */
void fm_read(const mxArray* a, double* out) {
  memcpy(out, a->data, numel(a) * (a->cplx ? 2 : 1) * sizeof(double));
}

long long fm_leaked(void) { return leaked; }
int fm_locked(void) { return lock_count; }

/* MATLAB clearing the MEX file / quitting: the exit handler runs once */
int fm_exit(void) {
  void (*fn)(void) = exit_fn;
  exit_fn = NULL;
  lock_count = 0;
  if (!fn) return 0;
  fn();
  return 1;
}
