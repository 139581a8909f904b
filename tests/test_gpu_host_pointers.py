"""Host-pointer mode through the C ABI: numpy arrays passed straight to the
entry points (as a MEX gateway would pass mxArray data) give bitwise the
same results as device arrays."""
import ctypes

import numpy as np
import pytest
import torch

from fasttools2d_amd import _capi
from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W

pytestmark = pytest.mark.gpu

D = ctypes.c_void_p


def _h(a):
    return D(a.ctypes.data) if a.size else D(0)


def test_host_pointer_calls_match_device_calls(gpu):
    pe = gpu
    w = W.uniform_box(1500, seed=12, targets_equal_sources=False, Nt=700)
    p = Pm.select_params("dlp", 1500, 700, w.f, 1.0, 1.0, tol=1e-10, M=64)
    ps, pt, f, nv = (np.ascontiguousarray(a) for a in (w.psrc, w.ptar, w.f, w.n))
    du = np.random.default_rng(3).standard_normal((700, 2))

    # device-pointer reference results through the Python mirror
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(a).to(dev)
    ref_r = pe.mex_stokes_dlp_real(t(ps), t(pt), t(f), t(nv), p.xi, p.nside_x, p.nside_y, 1.0, 1.0).cpu().numpy()
    ref_k = pe.mex_stokes_dlp_kspace(t(ps), t(pt), p.xi, p.eta, t(f), t(nv), p.Mx, p.My, 1.0, 1.0, p.w, p.P).cpu().numpy()
    xs = t(ps)
    ref_same = pe.mex_stokes_dlp_real(xs, xs, t(f), t(nv), p.xi, p.nside_x, p.nside_y, 1.0, 1.0).cpu().numpy()
    ref_g = {k: v.cpu().numpy() for k, v in pe.mex_stokes_dlp_kspace_backward(
        t(ps), t(pt), p.xi, p.eta, t(f), t(nv), p.Mx, p.My, 1.0, 1.0, p.w, p.P, t(du)).items()}
    torch.cuda.synchronize()

    lib = _capi.load()
    h = ctypes.c_void_p()
    assert lib.se2p_create(ctypes.byref(h), 0) == 0
    try:
        assert lib.se2p_set_host_pointers(h, 1) == 0
        u = np.zeros((700, 2))
        assert lib.se2p_dlp_real(h, _h(ps), 1500, _h(pt), 700, _h(f), _h(nv), p.xi, p.nside_x,
                                 p.nside_y, 1.0, 1.0, _h(u)) == 0
        assert np.array_equal(u, ref_r)
        u = np.zeros((700, 2))
        assert lib.se2p_dlp_kspace(h, _h(ps), 1500, _h(pt), 700, p.xi, p.eta, _h(f), _h(nv),
                                   p.Mx, p.My, 1.0, 1.0, p.w, p.P, _h(u)) == 0
        assert np.array_equal(u, ref_k)
        # the same host array as sources and targets: one staged copy, shared cell lists
        u = np.zeros((1500, 2))
        assert lib.se2p_dlp_real(h, _h(ps), 1500, _h(ps), 1500, _h(f), _h(nv), p.xi, p.nside_x,
                                 p.nside_y, 1.0, 1.0, _h(u)) == 0
        assert np.array_equal(u, ref_same)
        g = {k: np.zeros(v.shape) for k, v in ref_g.items()}
        assert lib.se2p_dlp_kspace_backward(h, _h(ps), 1500, _h(pt), 700, p.xi, p.eta, _h(f), _h(nv),
                                            p.Mx, p.My, 1.0, 1.0, p.w, p.P, _h(du), _h(g["psrc"]),
                                            _h(g["ptar"]), _h(g["f"]), _h(g["n"])) == 0
        for k in ref_g:
            assert np.array_equal(g[k], ref_g[k]), k
        # multi-box descriptors pointing at host arrays
        descs = (_capi.Se2pBoxDesc * 2)()
        outs = [np.zeros((700, 2)), np.zeros((1500, 2))]
        for i, (tt, nt) in enumerate(((pt, 700), (ps, 1500))):
            d = descs[i]
            d.psrc, d.ns, d.ptar, d.nt = ps.ctypes.data, 1500, tt.ctypes.data, nt
            d.f, d.n = f.ctypes.data, nv.ctypes.data
            d.Lx = d.Ly = 1.0
            d.xi, d.nside_x, d.nside_y = p.xi, p.nside_x, p.nside_y
            d.u = outs[i].ctypes.data
        assert lib.se2p_dlp_real_multibox(h, descs, 2) == 0
        assert np.array_equal(outs[0], ref_r) and np.array_equal(outs[1], ref_same)
        # the driver entry
        prm = _capi.Se2pParams()
        prm.P, prm.tol = 24, 1e-10
        u_host = np.zeros((700, 2))
        assert lib.se2p_dlp_ewald(h, _h(ps), 1500, _h(pt), 700, _h(f), _h(nv), 1.0, 1.0,
                                  ctypes.byref(prm), _h(u_host), D(0), D(0)) == 0
        assert np.all(np.isfinite(u_host)) and np.abs(u_host).max() > 0
    finally:
        lib.se2p_destroy(h)


def test_host_pointer_same_array_with_more_targets_than_sources(gpu):
    """psrc and ptar the same host array with Nt > Ns (the first Ns points are
    the sources, all Nt points the targets): the array is staged once, at the
    larger of the two sizes, and the results match device-pointer calls."""
    pe = gpu
    w = W.uniform_box(1200, seed=17)
    ns, nt = 500, 1200
    pts = np.ascontiguousarray(w.psrc)
    f, nv = np.ascontiguousarray(w.f[:ns]), np.ascontiguousarray(w.n[:ns])
    p = Pm.select_params("dlp", ns, nt, f, 1.0, 1.0, tol=1e-10, M=64)
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(a).to(dev)
    dp = t(pts)
    ref_r = pe.mex_stokes_dlp_real(dp[:ns], dp, t(f), t(nv), p.xi, p.nside_x, p.nside_y, 1.0, 1.0).cpu().numpy()
    ref_k = pe.mex_stokes_dlp_kspace(dp[:ns], dp, p.xi, p.eta, t(f), t(nv), p.Mx, p.My, 1.0, 1.0,
                                     p.w, p.P).cpu().numpy()
    torch.cuda.synchronize()
    lib = _capi.load()
    h = ctypes.c_void_p()
    assert lib.se2p_create(ctypes.byref(h), 0) == 0
    try:
        assert lib.se2p_set_host_pointers(h, 1) == 0
        u = np.zeros((nt, 2))
        assert lib.se2p_dlp_real(h, _h(pts), ns, _h(pts), nt, _h(f), _h(nv), p.xi, p.nside_x,
                                 p.nside_y, 1.0, 1.0, _h(u)) == 0, lib.se2p_last_error(h)
        assert np.array_equal(u, ref_r)
        u = np.zeros((nt, 2))
        assert lib.se2p_dlp_kspace(h, _h(pts), ns, _h(pts), nt, p.xi, p.eta, _h(f), _h(nv), p.Mx,
                                   p.My, 1.0, 1.0, p.w, p.P, _h(u)) == 0, lib.se2p_last_error(h)
        assert np.array_equal(u, ref_k)
    finally:
        lib.se2p_destroy(h)
