// Shared part of the MEX gateways over libstaqdic (include/staq_dic.h): one
// library context per MATLAB process, created on first use and destroyed by
// mexAtExit; fatal library errors become mexErrMsgTxt with the library's
// message.  Per-subset failures are never fatal: they come back as status
// codes, as the try/catch of LocalICGNQuadtree.m:92-97 makes them.
//
// Build a gateway with MATLAB's mex against the built library, e.g.
//   mex -output ba_interp2 ba_interp2_mex.cpp -I<repo>/include -L<repo>/staq_dic_amd -lstaqdic
#pragma once

#include <cstring>
#include <vector>

#include "mex.h"
#include "staq_dic.h"

namespace staq_mex {

inline staq_ctx *&ctx_slot()
{
    static staq_ctx *c = nullptr;
    return c;
}

inline void at_exit()
{
    staq_ctx_destroy(ctx_slot());
    ctx_slot() = nullptr;
}

inline staq_ctx *ctx()
{
    if (!ctx_slot()) {
        staq_ctx *c = nullptr;
        if (staq_ctx_create(&c, 0) != STAQ_OK) mexErrMsgTxt("staq: no usable GPU (staq_ctx_create failed)");
        ctx_slot() = c;
        mexAtExit(at_exit);
    }
    return ctx_slot();
}

inline void check(int rc)
{
    if (rc != STAQ_OK) mexErrMsgTxt(staq_last_error(ctx_slot()));
}

// a scalar or N-vector of window sizes as the library's per-node longs
inline std::vector<long> per_node(const mxArray *a, long N, const char *what)
{
    const long n = (long)mxGetNumberOfElements(a);
    if (!mxIsDouble(a) || (n != 1 && n != N)) mexErrMsgTxt(what);
    std::vector<long> out((size_t)N);
    const double *p = mxGetPr(a);
    for (long i = 0; i < N; ++i) out[(size_t)i] = (long)p[n == 1 ? 0 : i];
    return out;
}

inline double scalar(const mxArray *a, const char *what)
{
    if (!mxIsDouble(a) || mxGetNumberOfElements(a) != 1) mexErrMsgTxt(what);
    return mxGetScalar(a);
}

// the images of one frame pair (f = fNormalized .* mask, g, mask or [] for ones)
inline void set_frame(staq_ctx *c, const mxArray *f, const mxArray *g, const mxArray *mask)
{
    const long H = (long)mxGetM(f), W = (long)mxGetN(f);
    if (!mxIsDouble(f) || !mxIsDouble(g) || (long)mxGetM(g) != H || (long)mxGetN(g) != W)
        mexErrMsgTxt("f and g must be double arrays of the same size");
    const double *m = nullptr;
    if (mask && mxGetNumberOfElements(mask) > 0) {
        if (!mxIsDouble(mask) || (long)mxGetM(mask) != H || (long)mxGetN(mask) != W)
            mexErrMsgTxt("mask must be a double array of the image size (or [])");
        m = mxGetPr(mask);
    }
    check(staq_set_frame(c, 0, mxGetPr(f), mxGetPr(g), m, H, W));
}

}  // namespace staq_mex
