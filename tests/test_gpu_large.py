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


def test_adjoint_finite_differences_4e6_points(gpu, big):
    """Loss 1/2 |u - u*|^2 on the full Stokeslet Ewald sum (real + Fourier) at
    N = 4e6, with u* the sum at positions perturbed by N(0, (1e-4)^2) and
    densities scaled by 1 + 1e-3 N(0,1).  The backward entries' position
    gradient (dpsrc + dptar: targets are the sources) is checked against
    central differences on 16 points (seed 55), step 1e-7, and the density
    gradient by the exact linearity of the loss's first variation in f."""
    w, _ = big
    p = Pm.select_params("slp", N, N, w.f, 1.0, 1.0, tol=1e-12, M=2048)
    rng = np.random.default_rng(55)

    def forward(x, f):
        xs, fs = _dev(x), _dev(f)
        ur = gpu.mex_stokes_slp_real(xs, xs, fs, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
        uk = gpu.mex_stokes_slp_kspace(xs, xs, p.xi, p.eta, fs, p.Mx, p.My, 1.0, 1.0, p.w, p.P)
        return (ur + uk).cpu().numpy()

    x0, f0 = w.psrc, w.f
    xstar = Pm.wrap(x0 + 1e-4 * rng.standard_normal(x0.shape), 1.0)
    fstar = f0 * (1 + 1e-3 * rng.standard_normal(f0.shape))
    ustar = forward(xstar, fstar)
    u0 = forward(x0, f0)
    du = u0 - ustar
    xs, fs, dd = _dev(x0), _dev(f0), _dev(du)
    gr = gpu.mex_stokes_slp_real_backward(xs, xs, fs, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, dd)
    gk = gpu.mex_stokes_slp_kspace_backward(xs, xs, p.xi, p.eta, fs, p.Mx, p.My, 1.0, 1.0, p.w, p.P, dd)
    gpos = (gr["psrc"] + gr["ptar"] + gk["psrc"] + gk["ptar"]).cpu().numpy()
    gf = (gr["f"] + gk["f"]).cpu().numpy()

    # density gradient: d/de L(f + e v) at e = 0 equals <du, u(v)> exactly
    v = rng.standard_normal(f0.shape)
    uv = forward(x0, v)
    lin = float(np.sum(du * uv))
    assert abs(np.sum(gf * v) - lin) <= 1e-10 * np.sum(np.abs(gf * v)), (np.sum(gf * v), lin)

    def stable(xp):
        cell = tuple(np.floor((xp + 0.5) * p.nside_x).astype(int))
        return cell, cref.find_closest_node(xp[0], xp[1], 1.0, 1.0, 1.0 / p.Mx, p.P)[:2]

    step = 1e-7
    checked = 0
    worst = 0.0
    floor = 1e-3 * float(np.sqrt(np.mean(gpos ** 2)))   # scale for near-zero components
    for i in rng.choice(N, 64, replace=False):
        if checked == 16:
            break
        if np.any(np.abs(x0[i]) > 0.5 - 1e-4):
            continue
        moves = [x0[i] + s * step * np.eye(2)[c] for c in range(2) for s in (1, -1)]
        if any(stable(m) != stable(x0[i]) for m in moves):
            continue
        for c in range(2):
            xp, xm = x0.copy(), x0.copy()
            xp[i, c] += step
            xm[i, c] -= step
            up, um = forward(xp, f0), forward(xm, f0)
            # L+ - L- = 1/2 sum (u+ - u-) . (u+ + u- - 2 u*)
            fd = 0.5 * np.sum((up - um) * (up + um - 2 * ustar)) / (2 * step)
            err = abs(fd - gpos[i, c]) / max(abs(gpos[i, c]), floor)
            worst = max(worst, err)
        checked += 1
    print(f"[N=4e6 adjoint FD] {checked} points, worst relative error {worst:.2e}", flush=True)
    assert checked >= 8
    assert worst < 1e-4, worst
