/* Test-only declarations of the part of MATLAB's MEX C API (the pre-R2018a
 * separate-complex interface) that the gateways under marie_amd/mex use, so
 * that they can be compiled and exercised without MATLAB by mx_emul.cpp.
 * A MATLAB host builds the gateways against its own mex.h instead. */
#ifndef MARIE_TEST_MEX_H
#define MARIE_TEST_MEX_H
#include <stddef.h>

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef enum { mxDOUBLE_CLASS = 6 } mxClassID;

#ifdef __cplusplus
extern "C" {
#endif
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
int mexAtExit(void (*fn)(void));
double* mxGetPr(const mxArray* a);
double* mxGetPi(const mxArray* a);
bool mxIsComplex(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsChar(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
double mxGetScalar(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize len);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray* mxCreateDoubleScalar(double v);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid, mxComplexity c);
#ifdef __cplusplus
}
#endif

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
#endif
