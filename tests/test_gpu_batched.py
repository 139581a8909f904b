"""Batched right-hand sides: every density of a batch against the oracle and
against the single-density entry point (same geometry)."""
import numpy as np
import pytest
import torch

from oracle import cref
from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W
from tests.conftest import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.mark.parametrize("nrhs", [1, 3, 4, 7, 64])
def test_batched_matches_oracle_and_single(gpu, nrhs):
    pe = gpu
    w = W.uniform_box(1000, seed=1)
    pr = Pm.select_params("slp", 1000, 1000, w.f, 1.0, 1.0, tol=1e-10, M=64)
    pd = Pm.select_params("dlp", 1000, 1000, w.f, 1.0, 1.0, tol=1e-10, M=64)
    rng = np.random.default_rng(100 + nrhs)
    F = rng.standard_normal((nrhs, 1000, 2))
    ps, pt, nv = _dev(w.psrc), _dev(w.ptar), _dev(w.n)
    print(f"[batched nrhs={nrhs}]", flush=True)
    ur = pe.stokes_slp_real_batched(ps, pt, _dev(F), pr.xi, pr.nside_x, pr.nside_y, 1.0, 1.0).cpu().numpy()
    uk = pe.stokes_slp_kspace_batched(ps, pt, pr.xi, pr.eta, _dev(F), pr.Mx, pr.My, 1.0, 1.0, pr.w, pr.P).cpu().numpy()
    dr = pe.stokes_dlp_real_batched(ps, pt, _dev(F), nv, pd.xi, pd.nside_x, pd.nside_y, 1.0, 1.0).cpu().numpy()
    dk = pe.stokes_dlp_kspace_batched(ps, pt, pd.xi, pd.eta, _dev(F), nv, pd.Mx, pd.My, 1.0, 1.0, pd.w, pd.P).cpu().numpy()
    for r in sorted({0, nrhs // 2, nrhs - 1}):
        c = cref.slp_real(w.psrc, w.ptar, F[r], pr.xi, pr.nside_x, pr.nside_y, 1.0, 1.0)
        assert relerr(ur[r], c) <= TOL, ("slp_real", r)
        c = cref.slp_kspace(w.psrc, w.ptar, pr.xi, pr.eta, F[r], pr.Mx, pr.My, 1.0, 1.0, pr.w, pr.P)
        assert relerr(uk[r], c) <= TOL, ("slp_kspace", r)
        c = cref.dlp_real(w.psrc, w.ptar, F[r], w.n, pd.xi, pd.nside_x, pd.nside_y, 1.0, 1.0)
        assert relerr(dr[r], c) <= TOL, ("dlp_real", r)
        c = cref.dlp_kspace(w.psrc, w.ptar, pd.xi, pd.eta, F[r], w.n, pd.Mx, pd.My, 1.0, 1.0, pd.w, pd.P)
        assert relerr(dk[r], c) <= TOL, ("dlp_kspace", r)
        s = pe.mex_stokes_slp_real(ps, pt, _dev(F[r]), pr.xi, pr.nside_x, pr.nside_y, 1.0, 1.0).cpu().numpy()
        assert relerr(ur[r], s) <= 1e-14
        s = pe.mex_stokes_dlp_kspace(ps, pt, pd.xi, pd.eta, _dev(F[r]), nv, pd.Mx, pd.My, 1.0, 1.0, pd.w, pd.P).cpu().numpy()
        assert relerr(dk[r], s) <= 1e-14


def test_batched_ellipses_subset(gpu):
    """Stresslet workload geometry with 8 densities, checked on 128 targets."""
    pe = gpu
    w, p = W.dlp_ellipses_config()
    rng = np.random.default_rng(7)
    F = w.f[None, :, :] * (1 + 0.1 * rng.standard_normal((8, len(w.f), 2)))
    ps, nv = _dev(w.psrc), _dev(w.n)
    print("[batched ellipses nrhs=8]", flush=True)
    dr = pe.stokes_dlp_real_batched(ps, ps, _dev(F), nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0).cpu().numpy()
    dk = pe.stokes_dlp_kspace_batched(ps, ps, p.xi, p.eta, _dev(F), nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P).cpu().numpy()
    sub = np.random.default_rng(3).choice(len(w.psrc), 128, replace=False)
    for r in (0, 5):
        c = cref.dlp_real(w.psrc, w.ptar[sub], F[r], w.n, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
        assert relerr(dr[r][sub], c) <= TOL
        c = cref.dlp_kspace(w.psrc, w.ptar[sub], p.xi, p.eta, F[r], w.n, p.Mx, p.My, 1.0, 1.0, p.w, p.P)
        assert relerr(dk[r][sub], c) <= TOL
