// Z = ba_interp2(F, X, Y [, method]) — drop-in gateway for the reference's
// MEX (ba_interp2.cpp:259-335) over staq_ba_interp2 / staq_ba_interp2_f32.
// Same argument rules, error texts and output shape: X's dimensions followed
// by F's dimensions beyond the second; F, X, Y all double or all single, the
// output of that class.  The method argument is accepted and, as in the
// reference, every call interpolates with the Keys cubic.
#include "staq_mex.h"

extern "C" void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 3 && nrhs != 4)
        mexErrMsgTxt("Wrong number of input arguments for Z = ba_interp2(F, X, Y, [method])");
    if (nlhs > 1)
        mexErrMsgTxt("Wrong number of output arguments for Z = ba_interp2(F, X, Y, [method])");
    const bool dbl = mxIsDouble(prhs[0]) && mxIsDouble(prhs[1]) && mxIsDouble(prhs[2]);
    const bool sgl = mxIsSingle(prhs[0]) && mxIsSingle(prhs[1]) && mxIsSingle(prhs[2]);
    if (!dbl && !sgl)
        mexErrMsgTxt("ba_interp2 takes only double or single arguments for IMG,X,Y, and all must "
                     "have the same class");
    const mwSize ndf = mxGetNumberOfDimensions(prhs[0]);
    const mwSize ndx = mxGetNumberOfDimensions(prhs[1]);
    const mwSize *fd = mxGetDimensions(prhs[0]);
    const mwSize *xd = mxGetDimensions(prhs[1]);
    const mwSize *yd = mxGetDimensions(prhs[2]);
    if (ndx != mxGetNumberOfDimensions(prhs[2])) mexErrMsgTxt("X, Y should have the same size");
    if (ndx + ndf - 2 > 50) mexErrMsgTxt("Can't have that many dimensions in interpolated data.");
    std::vector<mwSize> od(ndx + ndf - 2);
    for (mwSize i = 0; i < ndx; ++i) {
        if (xd[i] != yd[i]) mexErrMsgTxt("X, Y should have the same size");
        od[i] = xd[i];
    }
    long P = 1;
    for (mwSize i = 2; i < ndf; ++i) {
        od[ndx + i - 2] = fd[i];
        P *= (long)fd[i];
    }
    plhs[0] = mxCreateNumericArray((mwSize)od.size(), od.data(), sgl ? mxSINGLE_CLASS : mxDOUBLE_CLASS,
                                   mxREAL);
    const long M = (long)fd[0], N = (long)fd[1], nd = (long)mxGetNumberOfElements(prhs[1]);
    if (nd == 0 || M == 0 || N == 0 || P == 0) return;
    staq_ctx *c = staq_mex::ctx();
    if (sgl)
        staq_mex::check(staq_ba_interp2_f32(c, (const float *)mxGetData(prhs[0]), M, N, P,
                                            (const float *)mxGetData(prhs[1]),
                                            (const float *)mxGetData(prhs[2]), nd,
                                            (float *)mxGetData(plhs[0])));
    else
        staq_mex::check(staq_ba_interp2(c, mxGetPr(prhs[0]), M, N, P, mxGetPr(prhs[1]),
                                        mxGetPr(prhs[2]), nd, mxGetPr(plhs[0])));
}
