// Implementation of the test stand-in MEX API (mex.h beside it) and a C
// driver for ctypes: build arrays from host buffers, call the gateway's
// mexFunction, read outputs back.  mexErrMsgTxt throws; the driver catches it
// and returns the message, as MATLAB would raise it.
#include "mex.h"

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct fake_mx_array {
    mxClassID cls = mxDOUBLE_CLASS;
    std::vector<mwSize> dims;
    std::vector<unsigned char> data;
};

namespace {
struct MexError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
std::vector<void (*)(void)> g_exit;

size_t elem_size(mxClassID c) { return c == mxSINGLE_CLASS ? 4 : 8; }

mxArray *make(mwSize ndim, const mwSize *dims, mxClassID cls)
{
    mxArray *a = new fake_mx_array();
    a->cls = cls;
    // MATLAB keeps at least two dimensions and drops trailing singletons beyond them
    a->dims.assign(dims, dims + ndim);
    while (a->dims.size() < 2) a->dims.push_back(1);
    while (a->dims.size() > 2 && a->dims.back() == 1) a->dims.pop_back();
    size_t n = 1;
    for (mwSize d : a->dims) n *= d;
    a->data.assign(n * elem_size(cls), 0);
    return a;
}
}  // namespace

extern "C" {

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity)
{
    const mwSize d[2] = {m, n};
    return make(2, d, mxDOUBLE_CLASS);
}
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity)
{
    if (cls != mxDOUBLE_CLASS && cls != mxSINGLE_CLASS) throw MexError("fake mx: class not supported");
    return make(ndim, dims, cls);
}
void mxDestroyArray(mxArray *a) { delete a; }
double *mxGetPr(const mxArray *a)
{
    return a->cls == mxDOUBLE_CLASS ? (double *)a->data.data() : nullptr;
}
void *mxGetData(const mxArray *a) { return (void *)a->data.data(); }
mwSize mxGetM(const mxArray *a) { return a->dims[0]; }
mwSize mxGetN(const mxArray *a)
{
    mwSize n = 1;
    for (size_t i = 1; i < a->dims.size(); ++i) n *= a->dims[i];
    return n;
}
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->dims.size(); }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims.data(); }
size_t mxGetNumberOfElements(const mxArray *a) { return a->data.size() / elem_size(a->cls); }
bool mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsSingle(const mxArray *a) { return a->cls == mxSINGLE_CLASS; }
mxClassID mxGetClassID(const mxArray *a) { return a->cls; }
double mxGetScalar(const mxArray *a)
{
    if (a->data.empty()) return 0.0;
    return a->cls == mxDOUBLE_CLASS ? *(const double *)a->data.data() : *(const float *)a->data.data();
}
void mexErrMsgTxt(const char *msg) { throw MexError(msg); }
int mexAtExit(void (*fn)(void))
{
    g_exit.push_back(fn);
    return 0;
}

// ---- driver for ctypes ---------------------------------------------------
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

// class: 6 double, 7 single; data (may be null) is copied in
mxArray *fakemx_array(int cls, int ndim, const long *dims, const void *data)
{
    std::vector<mwSize> d(dims, dims + ndim);
    mxArray *a = make((mwSize)ndim, d.data(), (mxClassID)cls);
    if (data) std::memcpy(a->data.data(), data, a->data.size());
    return a;
}
int fakemx_class(const mxArray *a) { return (int)a->cls; }
int fakemx_ndim(const mxArray *a) { return (int)a->dims.size(); }
long fakemx_dim(const mxArray *a, int k) { return (long)a->dims[(size_t)k]; }
const void *fakemx_data(const mxArray *a) { return a->data.data(); }
long fakemx_numel(const mxArray *a) { return (long)mxGetNumberOfElements(a); }
void fakemx_free(mxArray *a) { delete a; }

// 0: ok; 1: mexErrMsgTxt (message in err); 2: another exception
int fakemx_call(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *err, int errlen)
{
    try {
        mexFunction(nlhs, plhs, nrhs, prhs);
        return 0;
    } catch (const MexError &e) {
        std::strncpy(err, e.what(), (size_t)errlen - 1);
        err[errlen - 1] = 0;
        return 1;
    } catch (const std::exception &e) {
        std::strncpy(err, e.what(), (size_t)errlen - 1);
        err[errlen - 1] = 0;
        return 2;
    }
}

// what MATLAB does when the MEX file is cleared
void fakemx_clear()
{
    for (auto it = g_exit.rbegin(); it != g_exit.rend(); ++it) (*it)();
    g_exit.clear();
}

}  // extern "C"
