/* Test-only implementation of the fake mex.h: arrays on the C heap, errors recorded and unwound
 * with longjmp to the caller of call_mex(). */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

static jmp_buf g_jmp;
char g_err_id[128];
char g_err_msg[512];

static size_t elsize(mxClassID c) {
  switch (c) {
    case mxCHAR_CLASS: case mxUINT8_CLASS: return 1;
    case mxDOUBLE_CLASS: case mxUINT64_CLASS: return 8;
    default: return 4;
  }
}

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity cx) {
  (void)cx;
  size_t n = 1;
  for (mwSize d = 0; d < ndim; ++d) n *= dims[d];
  mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
  a->cls = cls;
  a->n = n;
  a->data = calloc(n ? n : 1, elsize(cls));
  return a;
}

void *mxGetData(const mxArray *a) { return a->data; }
size_t mxGetNumberOfElements(const mxArray *a) { return a->n; }

char *mxArrayToString(const mxArray *a) {
  if (a->cls != mxCHAR_CLASS) return NULL;
  char *s = (char *)malloc(a->n + 1);
  memcpy(s, a->data, a->n);
  s[a->n] = 0;
  return s;
}

void mxFree(void *p) { free(p); }

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  snprintf(g_err_id, sizeof g_err_id, "%s", id);
  vsnprintf(g_err_msg, sizeof g_err_msg, fmt, ap);
  va_end(ap);
  longjmp(g_jmp, 1);
}

int mexPrintf(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  int r = vprintf(fmt, ap);
  va_end(ap);
  return r;
}

/* Returns 0 on success, 1 when the gateway raised an error (id in g_err_id). */
int call_mex(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  g_err_id[0] = 0;
  if (setjmp(g_jmp)) return 1;
  mexFunction(nlhs, plhs, nrhs, prhs);
  return 0;
}
