"""The backward entries against the gradient of the FP64 direct Ewald sum.

L(x, f, n) = sum_t du_t . u_t with u the full periodic sum (real + Fourier
MEX outputs; the stresslet without its k = 0 mode, as the MEX kernels).  The
GPU gradient is the sum of the real- and Fourier-space backward entries
(dpsrc + dptar: the points are both).  The truth is the direct Ewald sum of
oracle/direct_sums.py (3x3-image real sum without cutoff plus the direct
Fourier sum over j in (-kinf, kinf]^2), which is independent of the split
and of everything the GPU path does (cells, cutoffs, gridding, FFTs):

* positions: central differences of L_direct, step 1e-7;
* densities and normals: L is linear in each, so <dL/df, v> = L_direct with
  density v (and likewise for n), exactly up to truncation.

Points are 60 uniform points at least 0.05 apart (periodic), targets = sources,
parameters from the driver formulas at tol 1e-13."""
import numpy as np
import pytest
import torch

from fasttools2d_amd import params as Pm
from oracle import direct_sums as ds

pytestmark = pytest.mark.gpu

N = 60


def _points(rng):
    pts = []
    while len(pts) < N:
        c = rng.uniform(-0.5, 0.5, 2)
        ok = True
        for q in pts:
            d = c - q
            d -= np.round(d)
            if np.hypot(*d) < 0.05:
                ok = False
                break
        if ok:
            pts.append(c)
    return np.array(pts)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_gradient_matches_direct_ewald(gpu, kind):
    pe = gpu
    rng = np.random.default_rng(91 if kind == "slp" else 92)
    x = _points(rng)
    f = rng.standard_normal((N, 2))
    ang = rng.uniform(0, 2 * np.pi, N)
    n = np.stack([np.cos(ang), np.sin(ang)], 1)
    du = rng.standard_normal((N, 2))
    p = Pm.select_params(kind, N, N, f, 1.0, 1.0, tol=1e-13)
    kinf = int(np.ceil(3.0 * p.xi))
    print(f"[adjoint vs direct {kind}] nside {p.nside_x} xi {p.xi:.3f} M {p.Mx} kinf {kinf}",
          flush=True)

    def u_direct(xx, ff, nn):
        if kind == "slp":
            return (ds.slp_real_ds(xx, xx, ff, 1.0, 1.0, p.xi) +
                    ds.slp_kspace_ds(xx, xx, ff, 1.0, 1.0, p.xi, kinf))
        return (ds.dlp_real_ds(xx, xx, ff, nn, 1.0, 1.0, p.xi) +
                ds.dlp_kspace_ds(xx, xx, ff, nn, 1.0, 1.0, p.xi, kinf, zero_mode=False))

    def L(xx, ff=f, nn=n):
        return float(np.sum(du * u_direct(xx, ff, nn)))

    xs, fs, ns_, dd = _t(x), _t(f), _t(n), _t(du)
    if kind == "slp":
        ur = pe.mex_stokes_slp_real(xs, xs, fs, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
        uk = pe.mex_stokes_slp_kspace(xs, xs, p.xi, p.eta, fs, p.Mx, p.My, 1.0, 1.0, p.w, p.P)
        gr = pe.mex_stokes_slp_real_backward(xs, xs, fs, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, dd)
        gk = pe.mex_stokes_slp_kspace_backward(xs, xs, p.xi, p.eta, fs, p.Mx, p.My, 1.0, 1.0,
                                               p.w, p.P, dd)
    else:
        ur = pe.mex_stokes_dlp_real(xs, xs, fs, ns_, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
        uk = pe.mex_stokes_dlp_kspace(xs, xs, p.xi, p.eta, fs, ns_, p.Mx, p.My, 1.0, 1.0, p.w,
                                      p.P)
        gr = pe.mex_stokes_dlp_real_backward(xs, xs, fs, ns_, p.xi, p.nside_x, p.nside_y, 1.0,
                                             1.0, dd)
        gk = pe.mex_stokes_dlp_kspace_backward(xs, xs, p.xi, p.eta, fs, ns_, p.Mx, p.My, 1.0,
                                               1.0, p.w, p.P, dd)
    u = (ur + uk).cpu().numpy()
    ud = u_direct(x, f, n)
    eu = np.abs(u - ud).max() / np.abs(ud).max()
    gx = (gr["psrc"] + gr["ptar"] + gk["psrc"] + gk["ptar"]).cpu().numpy()
    gf = (gr["f"] + gk["f"]).cpu().numpy()

    # positions: central differences of the direct sum
    h = 1e-7
    scale = np.abs(gx).max()
    worst = 0.0
    for i in rng.choice(N, 6, replace=False):
        for c in range(2):
            xp, xm = x.copy(), x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            fd = (L(xp) - L(xm)) / (2 * h)
            worst = max(worst, abs(fd - gx[i, c]) / scale)
    # densities (and normals): exact linearity of L
    v = rng.standard_normal((N, 2))
    ef = abs(np.sum(gf * v) - L(x, ff=v)) / np.sum(np.abs(gf * v))
    msg = f"  forward {eu:.2e}, positions (FD) {worst:.2e}, densities {ef:.2e}"
    if kind == "dlp":
        gn = (gr["n"] + gk["n"]).cpu().numpy()
        wv = rng.standard_normal((N, 2))
        en = abs(np.sum(gn * wv) - L(x, nn=wv)) / np.sum(np.abs(gn * wv))
        msg += f", normals {en:.2e}"
        assert en <= 1e-10
    print(msg, flush=True)
    assert eu <= 1e-11
    assert worst <= 1e-6
    assert ef <= 1e-10
