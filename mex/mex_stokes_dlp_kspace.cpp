// mex_stokes_dlp_kspace.cpp -- MEX gateway: uk = mex_stokes_dlp_kspace(psrc, ptar, xi,
// eta, f, n, Mx, My, Lx, Ly, w, P) on the GPU through se2p_dlp_kspace.
// Same arguments and checks as PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_kspace.cpp:
// 10-53 (count and shapes :12-26); se2p_dlp_kspace validates the grid, w and P.
#include "se2p_mex_common.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs != 12) mexErrMsgTxt("Incorrect number of input parameters");
    if (mxGetM(prhs[0]) != 2) mexErrMsgTxt("psrc must be a 2xn matrix.");
    if (mxGetM(prhs[1]) != 2) mexErrMsgTxt("ptar must be a 2xn matrix.");
    if (mxGetM(prhs[4]) != 2) mexErrMsgTxt("f must be a 2xn matrix.");
    if (mxGetM(prhs[5]) != 2) mexErrMsgTxt("n must be a 2xn matrix.");
    if (mxGetN(prhs[4]) != mxGetN(prhs[0])) mexErrMsgTxt("psrc and f must be the same size.");
    if (mxGetN(prhs[5]) != mxGetN(prhs[0])) mexErrMsgTxt("psrc and n must be the same size.");
    se2p_ctx *ctx = se2p_mex_context();
    const int64_t ns = static_cast<int64_t>(mxGetN(prhs[0]));
    const int64_t nt = static_cast<int64_t>(mxGetN(prhs[1]));
    double *u = se2p_mex_output(&plhs[0], nt);
    se2p_mex_check(se2p_dlp_kspace(ctx, mxGetPr(prhs[0]), ns, mxGetPr(prhs[1]), nt,
                                   mxGetScalar(prhs[2]), mxGetScalar(prhs[3]), mxGetPr(prhs[4]),
                                   mxGetPr(prhs[5]), se2p_mex_int(prhs[6]), se2p_mex_int(prhs[7]),
                                   mxGetScalar(prhs[8]), mxGetScalar(prhs[9]),
                                   mxGetScalar(prhs[10]), se2p_mex_int(prhs[11]), u));
}
