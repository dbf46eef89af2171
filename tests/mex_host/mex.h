/* Test-only stand-in for MATLAB's mex.h: just the mx/mex subset the simFDTD gateway uses.  It lets
 * the C gateway be compiled and driven without MATLAB (tests/mex_host/fake_mx.c implements it). */
#ifndef FAKE_MEX_H
#define FAKE_MEX_H
#include <stddef.h>
#include <stdint.h>

typedef size_t mwSize;
typedef enum { mxUNKNOWN_CLASS = 0, mxCHAR_CLASS, mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT32_CLASS, mxUINT32_CLASS,
               mxUINT64_CLASS, mxUINT8_CLASS } mxClassID;
typedef enum { mxREAL = 0 } mxComplexity;

typedef struct mxArray {
  mxClassID cls;
  size_t n;
  void *data;
} mxArray;

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity cx);
void *mxGetData(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
char *mxArrayToString(const mxArray *a);
void mxFree(void *p);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexPrintf(const char *fmt, ...);
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
#endif
