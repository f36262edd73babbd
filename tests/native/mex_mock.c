/* Test harness for matlab/ins2d_mex.c: a small implementation of the MEX API
 * of mex_mock/mex.h, with entry points that a test drives through ctypes:
 * build argument arrays, call the gateway (a MATLAB error comes back as a
 * nonzero return and its identifier and message), read the outputs. */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
  size_t n;
  double* pr;
  char* str;
  mxArray** cells;
  size_t ncells;
};

static jmp_buf g_jmp;
static char g_err[1024];
static void (*g_atexit)(void) = NULL;

double mxGetScalar(const mxArray* a) { return a->n ? a->pr[0] : 0.0; }
double* mxGetPr(const mxArray* a) { return a->pr; }
size_t mxGetNumberOfElements(const mxArray* a) { return a->cells ? a->ncells : (a->str ? strlen(a->str) : a->n); }
int mxGetString(const mxArray* a, char* buf, mwSize buflen) {
  if (!a->str || strlen(a->str) + 1 > buflen) return 1;
  strcpy(buf, a->str);
  return 0;
}
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
  (void)c;
  mxArray* a = calloc(1, sizeof *a);
  a->n = m * n;
  a->pr = calloc(a->n ? a->n : 1, sizeof(double));
  return a;
}
mxArray* mxCreateDoubleScalar(double v) {
  mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
  a->pr[0] = v;
  return a;
}
mxArray* mxGetCell(const mxArray* a, mwIndex i) { return i < a->ncells ? a->cells[i] : NULL; }
void mxDestroyArray(mxArray* a) {
  if (!a) return;
  free(a->pr);
  free(a->str);
  free(a->cells);
  free(a);
}
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  va_list ap;
  int k = snprintf(g_err, sizeof g_err, "%s: ", id);
  va_start(ap, fmt);
  vsnprintf(g_err + k, sizeof g_err - (size_t)k, fmt, ap);
  va_end(ap);
  longjmp(g_jmp, 1);
}
int mexAtExit(void (*fn)(void)) {
  g_atexit = fn;
  return 0;
}

/* ---- harness entry points (ctypes) */
mxArray* mock_double(const double* data, size_t n) {
  mxArray* a = mxCreateDoubleMatrix(n, 1, mxREAL);
  if (data) memcpy(a->pr, data, n * sizeof(double));
  return a;
}
mxArray* mock_string(const char* s) {
  mxArray* a = calloc(1, sizeof *a);
  a->str = malloc(strlen(s) + 1);
  strcpy(a->str, s);
  return a;
}
mxArray* mock_cell(mxArray* const* items, size_t n) {
  mxArray* a = calloc(1, sizeof *a);
  a->cells = calloc(n ? n : 1, sizeof(mxArray*));
  memcpy(a->cells, items, n * sizeof(mxArray*));
  a->ncells = n;
  return a;
}
double* mock_data(const mxArray* a) { return a->pr; }
size_t mock_numel(const mxArray* a) { return a->n; }
void mock_free(mxArray* a) { mxDestroyArray(a); }
const char* mock_error(void) { return g_err; }
/* call the gateway; 0 on success, 1 when it raised a MATLAB error */
int mock_call(int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs) {
  g_err[0] = 0;
  if (setjmp(g_jmp)) return 1;
  mexFunction(nlhs, plhs, nrhs, prhs);
  return 0;
}
/* what MATLAB does when the MEX file is cleared */
void mock_clear(void) {
  if (g_atexit) g_atexit();
}
