"""The MEX gateway sources (mex/*.cpp) against the C ABI, without MATLAB.

MATLAB's mex.h is not available here, so the gateways cannot be built.  This
test only checks that they are well-formed C++ calling include/se2p.h with
the right argument counts and types: it compiles each with -fsyntax-only
against declarations of the handful of MATLAB API functions they use
(written into a temporary directory here; nothing is linked or run), and it
checks that each gateway repeats its reference's argument checks."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GATEWAYS = {
    # name: (argument count, error messages the reference checks)
    "mex_stokes_slp_real": (8, []),
    "mex_stokes_slp_kspace": (11, ["psrc must be a 2xn matrix.", "ptar must be a 2xn matrix.",
                                   "f must be a 2xn matrix.", "psrc and f must be the same size."]),
    "mex_stokes_dlp_real": (9, ["Incorrect number of input parameters", "psrc must be a 2xn matrix.",
                                "ptar must be a 2xn matrix.", "f must be a 2xn matrix.",
                                "n must be a 2xn matrix.", "psrc and f must be the same size.",
                                "psrc and n must be the same size."]),
    "mex_stokes_dlp_kspace": (12, ["Incorrect number of input parameters",
                                   "psrc must be a 2xn matrix.", "ptar must be a 2xn matrix.",
                                   "f must be a 2xn matrix.", "n must be a 2xn matrix.",
                                   "psrc and f must be the same size.",
                                   "psrc and n must be the same size."]),
}

MX_DECLARATIONS = """
#pragma once
#include <cstddef>
typedef struct mxArray_tag mxArray;
typedef std::size_t mwSize;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;
mwSize mxGetM(const mxArray *);
mwSize mxGetN(const mxArray *);
double *mxGetPr(const mxArray *);
double mxGetScalar(const mxArray *);
mxArray *mxCreateDoubleMatrix(mwSize, mwSize, mxComplexity);
void mexErrMsgTxt(const char *);
int mexAtExit(void (*)(void));
"""


@pytest.mark.parametrize("name", sorted(GATEWAYS))
def test_gateway_compiles_against_the_c_abi(tmp_path, name):
    if shutil.which("g++") is None:
        pytest.skip("no C++ compiler")
    (tmp_path / "mex.h").write_text(MX_DECLARATIONS)
    src = os.path.join(ROOT, "mex", name + ".cpp")
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", str(tmp_path), "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "mex"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("name", sorted(GATEWAYS))
def test_gateway_repeats_the_reference_checks(name):
    nargs, messages = GATEWAYS[name]
    txt = open(os.path.join(ROOT, "mex", name + ".cpp")).read()
    assert f"if (nrhs != {nargs})" in txt
    for m in messages:
        assert f'"{m}"' in txt, m
    # the entry point it forwards to, and every prhs slot read
    assert re.search(r"se2p_" + name[len("mex_stokes_"):] + r"\(ctx,", txt)
    used = {int(k) for k in re.findall(r"prhs\[(\d+)\]", txt)}
    assert used == set(range(nargs))
