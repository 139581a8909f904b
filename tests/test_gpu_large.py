"""Maximum-size case: N = 4e6 uniform points in the unit box (targets =
sources, one array), tol 1e-12 parameters from the driver formulas.  The
Stokeslet runs on a 2048^2 grid (eta = 0.93); the stresslet on the driver's
own 3060^2 grid (2048^2 is below its 2 k_inf at tol 1e-12 and would give
eta > 1).  The GPU evaluates every target; 256 of them are compared with the
oracle at 1e-12."""
import numpy as np
import pytest
import torch

from oracle import cref
from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W
from tests.conftest import relerr

pytestmark = pytest.mark.gpu

N = 4_000_000


@pytest.fixture(scope="module")
def big():
    w = W.uniform_box(N, seed=5)
    sub = np.random.default_rng(5).choice(N, 256, replace=False)
    return w, sub


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def test_stokeslet_4e6_points_grid_2048(gpu, big):
    w, sub = big
    p = Pm.select_params("slp", N, N, w.f, 1.0, 1.0, tol=1e-12, M=2048)
    print(f"[N=4e6 slp] nside={p.nside_x} xi={p.xi:.4f} M={p.Mx} eta={p.eta:.3f}", flush=True)
    ps, f = _dev(w.psrc), _dev(w.f)
    ur = gpu.mex_stokes_slp_real(ps, ps, f, p.xi, p.nside_x, p.nside_y, 1.0, 1.0).cpu().numpy()
    uk = gpu.mex_stokes_slp_kspace(ps, ps, p.xi, p.eta, f, p.Mx, p.My, 1.0, 1.0, p.w, p.P).cpu().numpy()
    assert np.all(np.isfinite(ur)) and np.all(np.isfinite(uk))
    cr = cref.slp_real(w.psrc, w.ptar[sub], w.f, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
    ck = cref.slp_kspace(w.psrc, w.ptar[sub], p.xi, p.eta, w.f, p.Mx, p.My, 1.0, 1.0, p.w, p.P)
    er, ek = relerr(ur[sub], cr), relerr(uk[sub], ck)
    print(f"[N=4e6 slp] real {er:.2e} kspace {ek:.2e}", flush=True)
    assert er <= 1e-12 and ek <= 1e-12


def test_stresslet_4e6_points_driver_grid(gpu, big):
    w, sub = big
    p = Pm.select_params("dlp", N, N, w.f, 1.0, 1.0, tol=1e-12)
    print(f"[N=4e6 dlp] nside={p.nside_x} xi={p.xi:.4f} M={p.Mx} eta={p.eta:.3f}", flush=True)
    ps, f, nv = _dev(w.psrc), _dev(w.f), _dev(w.n)
    ur = gpu.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0).cpu().numpy()
    uk = gpu.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P).cpu().numpy()
    assert np.all(np.isfinite(ur)) and np.all(np.isfinite(uk))
    cr = cref.dlp_real(w.psrc, w.ptar[sub], w.f, w.n, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
    ck = cref.dlp_kspace(w.psrc, w.ptar[sub], p.xi, p.eta, w.f, w.n, p.Mx, p.My, 1.0, 1.0, p.w, p.P)
    er, ek = relerr(ur[sub], cr), relerr(uk[sub], ck)
    print(f"[N=4e6 dlp] real {er:.2e} kspace {ek:.2e}", flush=True)
    assert er <= 1e-12 and ek <= 1e-12
