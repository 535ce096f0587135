/*
 * mex.h -- test-only stand-in for MATLAB's external-interface header, enough
 * of it to compile and drive voxhenry_amd/mex/vh_mex.c without MATLAB.  The
 * names and signatures are MATLAB's documented C API (interleaved complex,
 * -R2018a); the behaviour is implemented in fake_mex.c, and fake_mex.c's fm_*
 * functions are the harness the tests call through ctypes.
 */
#ifndef FAKE_MEX_H
#define FAKE_MEX_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef double mxDouble;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct {
  mxDouble real, imag;
} mxComplexDouble;

size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsComplex(const mxArray* a);
bool mxIsSparse(const mxArray* a);
bool mxIsChar(const mxArray* a);
bool mxIsEmpty(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize buflen);
double mxGetScalar(const mxArray* a);
mxDouble* mxGetDoubles(const mxArray* a);
mxComplexDouble* mxGetComplexDoubles(const mxArray* a);
mwIndex* mxGetIr(const mxArray* a);
mwIndex* mxGetJc(const mxArray* a);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray* mxCreateDoubleScalar(double v);
void mxDestroyArray(mxArray* a);
void* mxMalloc(mwSize n);
void* mxCalloc(mwSize n, mwSize size);
void mxFree(void* p);

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
int mexPrintf(const char* fmt, ...);
int mexAtExit(void (*fn)(void));
void mexLock(void);
void mexUnlock(void);
bool mexIsLocked(void);

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#ifdef __cplusplus
}
#endif
#endif
