// [U, F, Hpar, status] = staq_local_icgn_mex(U0, coordinatesFEM, winsize, f, g, mask, tol [, plain])
//
// Gateway for the node loop of LocalICGNQuadtree.m:84-142 (funICGNQuadtree per
// node; with plain = 1 the loop of LocalICGN.m:60-115 and funICGN) over
// staq_local_icgn.  U0 is 2N x 1, coordinatesFEM N x 2 (MATLAB's own layout,
// passed through), winsize a scalar or N-vector, f = fNormalized .* mask, g,
// mask (or [] for ones) H x W, tol the IC-GN threshold.  Outputs: U 2N x 1,
// F 4N x 1, Hpar N x 21 and status N x 1 (the reference's ConvItPerEle codes:
// iterations 1..100, 101, 102, 103, -1), before the bad-subset
// classification and refill that stay in the .m file.
#include "staq_mex.h"

extern "C" void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 7 && nrhs != 8)
        mexErrMsgTxt("staq_local_icgn_mex(U0, coordinatesFEM, winsize, f, g, mask, tol [, plain]): "
                     "7 or 8 inputs expected");
    if (nlhs > 4) mexErrMsgTxt("staq_local_icgn_mex: at most 4 outputs (U, F, Hpar, status)");
    const mxArray *U0 = prhs[0], *coords = prhs[1];
    const long N = (long)mxGetM(coords);
    if (!mxIsDouble(coords) || mxGetN(coords) != 2 || N < 1)
        mexErrMsgTxt("coordinatesFEM must be an N x 2 double array");
    if (!mxIsDouble(U0) || (long)mxGetNumberOfElements(U0) != 2 * N)
        mexErrMsgTxt("U0 must hold 2N doubles [u1 v1 u2 v2 ...]");
    const std::vector<long> w = staq_mex::per_node(prhs[2], N, "winsize must be a scalar or N-vector");
    staq_params p;
    staq_params_default(&p);
    p.tol = staq_mex::scalar(prhs[6], "tol must be a double scalar");
    if (nrhs == 8) p.plain = staq_mex::scalar(prhs[7], "plain must be a double scalar") != 0;
    staq_ctx *c = staq_mex::ctx();
    staq_mex::set_frame(c, prhs[3], prhs[4], prhs[5]);
    plhs[0] = mxCreateDoubleMatrix(2 * N, 1, mxREAL);
    mxArray *F = mxCreateDoubleMatrix(4 * N, 1, mxREAL);
    mxArray *Hp = mxCreateDoubleMatrix(N, 21, mxREAL);
    mxArray *S = mxCreateDoubleMatrix(N, 1, mxREAL);
    std::vector<int> st((size_t)N);
    staq_mex::check(staq_local_icgn(c, &p, N, mxGetPr(U0), mxGetPr(coords), w.data(), nullptr,
                                    mxGetPr(plhs[0]), mxGetPr(F), mxGetPr(Hp), st.data()));
    double *sp = mxGetPr(S);
    for (long i = 0; i < N; ++i) sp[i] = st[(size_t)i];
    mxArray *outs[3] = {F, Hp, S};
    for (int k = 0; k < 3; ++k) {
        if (nlhs > k + 1)
            plhs[k + 1] = outs[k];
        else
            mxDestroyArray(outs[k]);
    }
}
