// se2p_mex_common.h -- shared by the four MEX gateways onto libse2p.so.
//
// Each MEX binary keeps one se2p context, created on its first call in
// host-pointer mode (mxArray data is host memory: the library stages inputs
// on the GPU and copies the output back) and destroyed when MATLAB clears the
// MEX file.  The GPU is device 0, or the one named by the environment
// variable SE2P_DEVICE.  Errors from the library abort the call through
// mexErrMsgTxt with the library's message, as the reference's own checks do.
#pragma once

#include <cstdint>
#include <cstdlib>

#include "mex.h"
#include "se2p.h"

static se2p_ctx *g_se2p_ctx = nullptr;

static void se2p_mex_cleanup()
{
    if (g_se2p_ctx) {
        se2p_destroy(g_se2p_ctx);
        g_se2p_ctx = nullptr;
    }
}

static se2p_ctx *se2p_mex_context()
{
    if (!g_se2p_ctx) {
        const char *dev = std::getenv("SE2P_DEVICE");
        if (se2p_create(&g_se2p_ctx, dev ? std::atoi(dev) : 0) != SE2P_OK) {
            g_se2p_ctx = nullptr;
            mexErrMsgTxt("se2p_create failed: no usable GPU");
        }
        se2p_set_host_pointers(g_se2p_ctx, 1);
        mexAtExit(se2p_mex_cleanup);
    }
    return g_se2p_ctx;
}

static void se2p_mex_check(int rc)
{
    if (rc != SE2P_OK) mexErrMsgTxt(se2p_last_error(g_se2p_ctx));
}

// Integer parameters the reference reads as static_cast<int>(mxGetScalar(.)).
static int se2p_mex_int(const mxArray *a) { return static_cast<int>(mxGetScalar(a)); }

// The 2 x Nt result, zero-filled by MATLAB and overwritten by the library.
static double *se2p_mex_output(mxArray **out, int64_t nt)
{
    *out = mxCreateDoubleMatrix(2, static_cast<size_t>(nt), mxREAL);
    return mxGetPr(*out);
}
