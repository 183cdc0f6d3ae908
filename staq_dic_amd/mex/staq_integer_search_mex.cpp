// [x0, y0, u, v, ccmax] = staq_integer_search_mex(f, g, r, gridx, gridy, winsize, winstepsize)
//
// Gateway for funIntegerSearch(f, g, r, gridx, gridy, winsize, winstepsize, 0)
// (func/funIntegerSearch.m:3-35, whole-field mode, :82-203) over
// staq_integer_search: f = ImgRef, g = ImgDef (H x W double), r a scalar or
// [r1 r2], gridx / gridy the ROI ranges [lo hi].  Outputs are ny x nx (rows =
// y, funIntegerSearch.m:108); ccmax is cc.max.  The hole masking and the
// outlier filter of IntegerSearch.m:69-119 stay in the .m file.
#include "staq_mex.h"

extern "C" void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 7)
        mexErrMsgTxt("staq_integer_search_mex(f, g, r, gridx, gridy, winsize, winstepsize): "
                     "7 inputs expected");
    if (nlhs > 5) mexErrMsgTxt("staq_integer_search_mex: at most 5 outputs (x0, y0, u, v, ccmax)");
    const mxArray *r = prhs[2], *gx = prhs[3], *gy = prhs[4];
    const long nr = (long)mxGetNumberOfElements(r);
    if (!mxIsDouble(r) || (nr != 1 && nr != 2)) mexErrMsgTxt("r must be a scalar or [r1 r2]");
    if (!mxIsDouble(gx) || mxGetNumberOfElements(gx) != 2 || !mxIsDouble(gy) ||
        mxGetNumberOfElements(gy) != 2)
        mexErrMsgTxt("gridx and gridy must be [lo hi]");
    const long w = (long)staq_mex::scalar(prhs[5], "winsize must be a double scalar");
    const long step = (long)staq_mex::scalar(prhs[6], "winstepsize must be a double scalar");
    if (w < 2 || step < 1) mexErrMsgTxt("winsize must be at least 2 and winstepsize at least 1");
    const long r1 = (long)mxGetPr(r)[0], r2 = (long)mxGetPr(r)[nr - 1];
    const long gxr[2] = {(long)mxGetPr(gx)[0], (long)mxGetPr(gx)[1]};
    const long gyr[2] = {(long)mxGetPr(gy)[0], (long)mxGetPr(gy)[1]};
    staq_ctx *c = staq_mex::ctx();
    staq_mex::set_frame(c, prhs[0], prhs[1], nullptr);
    const long H = (long)mxGetM(prhs[0]), W = (long)mxGetN(prhs[0]);
    const long mx = (H / step + 2) * (W / step + 2);
    std::vector<double> buf(5 * (size_t)mx);
    long nx = 0, ny = 0;
    staq_mex::check(staq_integer_search(c, 0, r1, r2, gxr, gyr, w, step, mx, &nx, &ny, buf.data(),
                                        buf.data() + mx, buf.data() + 2 * mx, buf.data() + 3 * mx,
                                        buf.data() + 4 * mx));
    for (int k = 0; k < 5 && k < (nlhs > 0 ? nlhs : 1); ++k) {
        plhs[k] = mxCreateDoubleMatrix(ny, nx, mxREAL);
        std::memcpy(mxGetPr(plhs[k]), buf.data() + (size_t)k * mx, sizeof(double) * nx * ny);
    }
}
