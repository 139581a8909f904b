"""The fused cell-list / bin sort against the multi-pass radix sort.

Both are stable sorts, so the sorted order -- and with it every summation
order downstream -- must be the same: the four sort modes (fused or radix
sort, run boxes from the final scatter or from their own kernel) give
bitwise identical results.  Cases cover the headline stresslet workload,
separate targets, several densities, adjoints, several boxes, a point set
that falls into one cell, sizes around a workgroup and the largest fused
size and the one past it.
"""
import numpy as np
import pytest
import torch

from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W

pytestmark = pytest.mark.gpu

MODES = (0, 1, 2, 3)   # 1: radix sort, 2: separate run-box kernel


@pytest.fixture(scope="module")
def contexts(gpu):
    pe = gpu
    ctxs = {}
    for m in MODES:
        c = pe.Context(0)
        c.set_sort_mode(m)
        ctxs[m] = c
    yield ctxs
    for c in ctxs.values():
        c.close()


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _all_equal(results, label):
    ref = results[0]
    for m, r in results.items():
        rr = r if isinstance(r, (list, tuple)) else [r]
        ref_l = ref if isinstance(ref, (list, tuple)) else [ref]
        for a, b in zip(rr, ref_l):
            assert torch.equal(a, b), f"{label}: mode {m} differs from mode 0"
    print(f"[{label}] {len(results)} modes bitwise equal", flush=True)


def _forward(pe, kind, ps, pt, f, nv, p, L, ctx):
    if kind == "slp":
        ur = pe.mex_stokes_slp_real(ps, pt, f, p.xi, p.nside_x, p.nside_y, L, L, ctx=ctx)
        uk = pe.mex_stokes_slp_kspace(ps, pt, p.xi, p.eta, f, p.Mx, p.My, L, L, p.w, p.P, ctx=ctx)
    else:
        ur = pe.mex_stokes_dlp_real(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, L, L, ctx=ctx)
        uk = pe.mex_stokes_dlp_kspace(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, L, L, p.w, p.P, ctx=ctx)
    torch.cuda.synchronize()
    return [ur, uk]


def test_headline_ellipses(gpu, contexts):
    w = W.ellipses(seed=2)
    n = w.psrc.shape[0]
    p = Pm.select_params("dlp", n, n, w.f, w.Lx, w.Ly, tol=1e-10, M=512)
    ps, f, nv = _t(w.psrc), _t(w.f), _t(w.n)
    res = {m: _forward(gpu, "dlp", ps, ps, f, nv, p, w.Lx, c) for m, c in contexts.items()}
    _all_equal(res, "dlp ellipses N=1e5, targets = sources")


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_separate_targets(gpu, contexts, kind):
    w = W.uniform_box(3000, seed=21, targets_equal_sources=False, Nt=1200)
    p = Pm.select_params(kind, 3000, 1200, w.f, 1.0, 1.0, tol=1e-10, M=64)
    ps, pt, f, nv = _t(w.psrc), _t(w.ptar), _t(w.f), _t(w.n)
    res = {m: _forward(gpu, kind, ps, pt, f, nv, p, 1.0, c) for m, c in contexts.items()}
    _all_equal(res, f"{kind} N=3000, 1200 separate targets")


@pytest.mark.parametrize("n", [37, 1024, 1025, 5000])
def test_small_sizes(gpu, contexts, n):
    w = W.uniform_box(n, seed=22 + n)
    p = Pm.select_params("dlp", n, n, w.f, 1.0, 1.0, tol=1e-10, M=32)
    ps, f, nv = _t(w.psrc), _t(w.f), _t(w.n)
    res = {m: _forward(gpu, "dlp", ps, ps, f, nv, p, 1.0, c) for m, c in contexts.items()}
    _all_equal(res, f"dlp N={n}")


@pytest.mark.parametrize("n", [229376, 229377])
def test_fused_size_limit(gpu, contexts, n):
    # 229376 points is the largest fused sort; one more takes the radix sort.
    # nside 29 keeps the cell keys (29^2 x 16 = 13456) within the fused sort.
    pe = gpu
    w = W.uniform_box(n, seed=5)
    p = Pm.select_params("slp", n, n, w.f, 1.0, 1.0, tol=1e-8, M=256)
    ps, f = _t(w.psrc), _t(w.f)
    res = {}
    for m in (0, 1):
        c = contexts[m]
        res[m] = [pe.mex_stokes_slp_real(ps, ps, f, p.xi, 29, 29, 1.0, 1.0, ctx=c),
                  pe.mex_stokes_slp_kspace(ps, ps, p.xi, p.eta, f, p.Mx, p.My, 1.0, 1.0, p.w, p.P,
                                           ctx=c)]
        torch.cuda.synchronize()
    _all_equal(res, f"slp N={n}, nside 29, grid {p.Mx}")


def test_clustered_points(gpu, contexts):
    # 6000 points inside one small square: one cell, a few long runs and a
    # single low-digit group that holds nearly every point
    rng = np.random.default_rng(8)
    ps_np = 0.1 + 2e-3 * rng.random((6000, 2))
    f_np = rng.standard_normal((6000, 2))
    ang = rng.uniform(0, 2 * np.pi, 6000)
    n_np = np.stack([np.cos(ang), np.sin(ang)], 1)
    p = Pm.select_params("dlp", 6000, 6000, f_np, 1.0, 1.0, tol=1e-10, M=64)
    ps, f, nv = _t(ps_np), _t(f_np), _t(n_np)
    res = {m: _forward(gpu, "dlp", ps, ps, f, nv, p, 1.0, c) for m, c in contexts.items()}
    _all_equal(res, "dlp, 6000 points in one cell")


def test_batched_and_backward(gpu, contexts):
    pe = gpu
    w = W.uniform_box(4000, seed=23)
    p = Pm.select_params("dlp", 4000, 4000, w.f, 1.0, 1.0, tol=1e-10, M=64)
    ps, nv = _t(w.psrc), _t(w.n)
    F = _t(np.random.default_rng(4).standard_normal((5, 4000, 2)))
    du = _t(np.random.default_rng(6).standard_normal((4000, 2)))
    res = {}
    for m, c in contexts.items():
        out = [pe.stokes_dlp_real_batched(ps, ps, F, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, ctx=c),
               pe.stokes_dlp_kspace_batched(ps, ps, p.xi, p.eta, F, nv, p.Mx, p.My, 1.0, 1.0, p.w,
                                            p.P, ctx=c)]
        g = pe.mex_stokes_dlp_real_backward(ps, ps, F[0], nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0,
                                            du, ctx=c)
        out += [g[k] for k in sorted(g)]
        torch.cuda.synchronize()
        res[m] = out
    _all_equal(res, "dlp batched (5 densities) and real backward")


def test_multibox(gpu, contexts):
    pe = gpu
    boxes = []
    for i, (n, L) in enumerate(((900, 1.0), (2500, 2.0), (1300, 0.7))):
        w = W.uniform_box(n, seed=30 + i, L=L)
        p = Pm.select_params("slp", n, n, w.f, L, L, tol=1e-10, M=32)
        boxes.append(dict(psrc=_t(w.psrc), ptar=_t(w.psrc), f=_t(w.f), Lx=L, Ly=L, xi=p.xi,
                          nside_x=p.nside_x, nside_y=p.nside_y, eta=p.eta, w=p.w, Mx=p.Mx, My=p.My,
                          P=p.P))
    res = {}
    for m, c in contexts.items():
        out = pe.stokes_slp_real_multibox(boxes, ctx=c) + pe.stokes_slp_kspace_multibox(boxes, ctx=c)
        torch.cuda.synchronize()
        res[m] = out
    _all_equal(res, "slp, 3 boxes")
