/* A minimal stand-in for MATLAB's MEX API (the functions matlab/ins2d_mex.c
 * calls, with MATLAB's documented signatures), so that the gateway can be
 * compiled and driven without MATLAB: the test harness in mex_mock.c creates
 * the arrays, calls mexFunction and reads the results. */
#ifndef INS2D_TEST_MEX_MOCK_H
#define INS2D_TEST_MEX_MOCK_H
#include <stddef.h>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

double mxGetScalar(const mxArray* a);
double* mxGetPr(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize buflen);
mxArray* mxCreateDoubleScalar(double v);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray* mxGetCell(const mxArray* a, mwIndex i);
void mxDestroyArray(mxArray* a);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
int mexAtExit(void (*fn)(void));

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
#endif
