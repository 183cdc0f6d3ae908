// [U, status] = staq_subpb1_mex(UOld, FOld, vdual, coordinatesFEM, winsize_List, f, g, mask, mu, tol [, plain])
//
// Gateway for the node loop of Subpb1Quadtree.m:84-159 (funICGN_Subpb1 per
// node; with plain = 1 the loop of Subpb1.m) over staq_subpb1.  UOld and
// vdual are 2N x 1, FOld 4N x 1, coordinatesFEM N x 2, winsize_List N x 2
// (wx, wy) or a scalar, mu the AL penalty.  Outputs: U 2N x 1 and status
// N x 1, where an error the reference catches is reported as 0, as
// Subpb1Quadtree records it in ConvItPerEle.
#include "staq_mex.h"

extern "C" void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 10 && nrhs != 11)
        mexErrMsgTxt("staq_subpb1_mex(UOld, FOld, vdual, coordinatesFEM, winsize_List, f, g, mask, "
                     "mu, tol [, plain]): 10 or 11 inputs expected");
    if (nlhs > 2) mexErrMsgTxt("staq_subpb1_mex: at most 2 outputs (U, status)");
    const mxArray *coords = prhs[3], *wl = prhs[4];
    const long N = (long)mxGetM(coords);
    if (!mxIsDouble(coords) || mxGetN(coords) != 2 || N < 1)
        mexErrMsgTxt("coordinatesFEM must be an N x 2 double array");
    if (!mxIsDouble(prhs[0]) || (long)mxGetNumberOfElements(prhs[0]) != 2 * N ||
        !mxIsDouble(prhs[2]) || (long)mxGetNumberOfElements(prhs[2]) != 2 * N)
        mexErrMsgTxt("UOld and vdual must hold 2N doubles");
    if (!mxIsDouble(prhs[1]) || (long)mxGetNumberOfElements(prhs[1]) != 4 * N)
        mexErrMsgTxt("FOld must hold 4N doubles");
    std::vector<long> wx((size_t)N), wy((size_t)N);
    const double *wp = mxIsDouble(wl) ? mxGetPr(wl) : nullptr;
    if (wp && mxGetNumberOfElements(wl) == 1) {
        for (long i = 0; i < N; ++i) wx[(size_t)i] = wy[(size_t)i] = (long)wp[0];
    } else if (wp && (long)mxGetM(wl) == N && mxGetN(wl) == 2) {
        for (long i = 0; i < N; ++i) {
            wx[(size_t)i] = (long)wp[i];
            wy[(size_t)i] = (long)wp[N + i];
        }
    } else {
        mexErrMsgTxt("winsize_List must be N x 2 (wx, wy) or a scalar");
    }
    staq_params p;
    staq_params_default(&p);
    p.mu = staq_mex::scalar(prhs[8], "mu must be a double scalar");
    p.tol = staq_mex::scalar(prhs[9], "tol must be a double scalar");
    if (nrhs == 11) p.plain = staq_mex::scalar(prhs[10], "plain must be a double scalar") != 0;
    staq_ctx *c = staq_mex::ctx();
    staq_mex::set_frame(c, prhs[5], prhs[6], prhs[7]);
    plhs[0] = mxCreateDoubleMatrix(2 * N, 1, mxREAL);
    std::vector<int> st((size_t)N);
    staq_mex::check(staq_subpb1(c, &p, N, mxGetPr(prhs[0]), mxGetPr(prhs[1]), nullptr,
                                mxGetPr(prhs[2]), mxGetPr(coords), wx.data(), wy.data(), nullptr,
                                mxGetPr(plhs[0]), st.data()));
    if (nlhs > 1) {
        plhs[1] = mxCreateDoubleMatrix(N, 1, mxREAL);
        double *sp = mxGetPr(plhs[1]);
        for (long i = 0; i < N; ++i) sp[i] = st[(size_t)i] == -1 ? 0.0 : st[(size_t)i];
    }
}
