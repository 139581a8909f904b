"""Accuracy of the HIP path against the FP64 direct Ewald sum (truth) on the
benchmark workload: 200 ellipses x 500 nodes (N = 1e5), grid 512^2, P = 24,
tol 1e-10.  The truth on 256 targets is the committed fixture
tests/golden/truth_dlp_ellipses_N1e5.json (3x3-image real-space sum plus
the direct Fourier sum over j in (-256, 256]^2), written by
tests/golden/make_truth_dlp_ellipses.py.  Reported per output (the
real-space and Fourier-space MEX outputs, their sum, per component)."""
import json
import os

import numpy as np
import pytest
import torch

from fasttools2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_dlp_ellipses_against_direct_ewald(gpu):
    pe = gpu
    path = os.path.join(os.path.dirname(__file__), "golden", "truth_dlp_ellipses_N1e5.json")
    with open(path) as fh:
        d = json.load(fh)
    w, p = W.dlp_ellipses_config(seed=2, n_ell=200, n_nodes=500, tol=1e-10, M=512)
    sub = np.array(d["targets"])
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ps, f, nv = t(w.psrc), t(w.f), t(w.n)
    ur = pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, w.Lx, w.Ly).cpu().numpy()
    uk = pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, w.Lx, w.Ly, p.w,
                                  p.P).cpu().numpy()
    tr, tk = np.array(d["u_real"]), np.array(d["u_kspace"])
    tt = tr + tk
    scale = np.abs(tt).max()
    u = ur[sub] + uk[sub]
    err = {
        "real": np.abs(ur[sub] - tr).max() / scale,
        "kspace": np.abs(uk[sub] - tk).max() / scale,
        "sum": np.abs(u - tt).max() / scale,
        "sum_u1": np.abs(u[:, 0] - tt[:, 0]).max() / np.abs(tt[:, 0]).max(),
        "sum_u2": np.abs(u[:, 1] - tt[:, 1]).max() / np.abs(tt[:, 1]).max(),
    }
    print("[dlp ellipses N=1e5 vs direct Ewald] " +
          ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    assert err["sum"] <= 1e-10
    assert err["sum_u1"] <= 1e-10 and err["sum_u2"] <= 1e-10
    assert err["kspace"] <= 1e-13
