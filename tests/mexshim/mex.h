// Test stand-in for MATLAB's MEX API: only the mx* / mex* routines the
// gateways under staq_dic_amd/mex use, with MATLAB's meaning (column-major
// arrays of at least two dimensions; mexErrMsgTxt does not return).  It lets
// the gateways build and run in tests without MATLAB; fake_mx.cpp implements
// it and adds a small C driver for ctypes.
#pragma once

#include <cstddef>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef enum { mxUNKNOWN_CLASS = 0, mxDOUBLE_CLASS = 6, mxSINGLE_CLASS = 7 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct fake_mx_array mxArray;

extern "C" {
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity c);
void mxDestroyArray(mxArray *a);
double *mxGetPr(const mxArray *a);
void *mxGetData(const mxArray *a);
mwSize mxGetM(const mxArray *a);
mwSize mxGetN(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
bool mxIsDouble(const mxArray *a);
bool mxIsSingle(const mxArray *a);
mxClassID mxGetClassID(const mxArray *a);
double mxGetScalar(const mxArray *a);
[[noreturn]] void mexErrMsgTxt(const char *msg);
int mexAtExit(void (*fn)(void));
}
