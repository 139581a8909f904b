// mex_stokes_slp_real.cpp -- MEX gateway: ur = mex_stokes_slp_real(psrc, ptar, f, xi,
// nside_x, nside_y, Lx, Ly) on the GPU through se2p_slp_real.
// Same arguments as PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_real.cpp:58-83,
// which checks none of them; the argument count is checked here so that a
// short call fails instead of reading past prhs, and se2p_slp_real validates
// the values (nside >= 1, positive xi and box).
#include "se2p_mex_common.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs != 8) mexErrMsgTxt("Incorrect number of input parameters");
    se2p_ctx *ctx = se2p_mex_context();
    const int64_t ns = static_cast<int64_t>(mxGetN(prhs[0]));
    const int64_t nt = static_cast<int64_t>(mxGetN(prhs[1]));
    double *u = se2p_mex_output(&plhs[0], nt);
    se2p_mex_check(se2p_slp_real(ctx, mxGetPr(prhs[0]), ns, mxGetPr(prhs[1]), nt, mxGetPr(prhs[2]),
                                 mxGetScalar(prhs[3]), se2p_mex_int(prhs[4]), se2p_mex_int(prhs[5]),
                                 mxGetScalar(prhs[6]), mxGetScalar(prhs[7]), u));
}
