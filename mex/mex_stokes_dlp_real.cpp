// mex_stokes_dlp_real.cpp -- MEX gateway: ur = mex_stokes_dlp_real(psrc, ptar, f, n, xi,
// nside_x, nside_y, Lx, Ly) on the GPU through se2p_dlp_real.
// Same arguments and checks as PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_real.cpp:
// 15-52 (count and shapes :20-34); se2p_dlp_real validates the values.
#include "se2p_mex_common.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs != 9) mexErrMsgTxt("Incorrect number of input parameters");
    if (mxGetM(prhs[0]) != 2) mexErrMsgTxt("psrc must be a 2xn matrix.");
    if (mxGetM(prhs[1]) != 2) mexErrMsgTxt("ptar must be a 2xn matrix.");
    if (mxGetM(prhs[2]) != 2) mexErrMsgTxt("f must be a 2xn matrix.");
    if (mxGetM(prhs[3]) != 2) mexErrMsgTxt("n must be a 2xn matrix.");
    if (mxGetN(prhs[2]) != mxGetN(prhs[0])) mexErrMsgTxt("psrc and f must be the same size.");
    if (mxGetN(prhs[3]) != mxGetN(prhs[0])) mexErrMsgTxt("psrc and n must be the same size.");
    se2p_ctx *ctx = se2p_mex_context();
    const int64_t ns = static_cast<int64_t>(mxGetN(prhs[0]));
    const int64_t nt = static_cast<int64_t>(mxGetN(prhs[1]));
    double *u = se2p_mex_output(&plhs[0], nt);
    se2p_mex_check(se2p_dlp_real(ctx, mxGetPr(prhs[0]), ns, mxGetPr(prhs[1]), nt, mxGetPr(prhs[2]),
                                 mxGetPr(prhs[3]), mxGetScalar(prhs[4]), se2p_mex_int(prhs[5]),
                                 se2p_mex_int(prhs[6]), mxGetScalar(prhs[7]),
                                 mxGetScalar(prhs[8]), u));
}
