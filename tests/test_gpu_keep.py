"""Kept forward state (se2p_set_keep): backward calls that reuse the forward
call's sorted state give bitwise the results of backward calls that rebuild
it in two passes (se2p_set_adjoint_passes(ctx, 2)), for all four sums, and
agree to 1e-13 with the default one-pass rebuild; a backward call whose
arguments differ from the kept forward call's (another density array) does
not reuse it; forward results are unchanged by keeping; the stresslet
benchmark (N = 1e5, grid 512^2) with kept state matches the oracle adjoints
on a target subset."""
import numpy as np
import pytest
import torch

from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W
from oracle import cref
from tests.conftest import relerr

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _fwd(pe, kind, ps, f, nv, p, Lx, Ly, ctx):
    if kind == "slp_real":
        return pe.mex_stokes_slp_real(ps, ps, f, p.xi, p.nside_x, p.nside_y, Lx, Ly, ctx=ctx)
    if kind == "dlp_real":
        return pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, Lx, Ly, ctx=ctx)
    if kind == "slp_kspace":
        return pe.mex_stokes_slp_kspace(ps, ps, p.xi, p.eta, f, p.Mx, p.My, Lx, Ly, p.w, p.P,
                                        ctx=ctx)
    return pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, Lx, Ly, p.w, p.P,
                                    ctx=ctx)


def _bwd(pe, kind, ps, f, nv, du, p, Lx, Ly, ctx):
    if kind == "slp_real":
        g = pe.mex_stokes_slp_real_backward(ps, ps, f, p.xi, p.nside_x, p.nside_y, Lx, Ly, du,
                                            ctx=ctx)
    elif kind == "dlp_real":
        g = pe.mex_stokes_dlp_real_backward(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, Lx, Ly, du,
                                            ctx=ctx)
    elif kind == "slp_kspace":
        g = pe.mex_stokes_slp_kspace_backward(ps, ps, p.xi, p.eta, f, p.Mx, p.My, Lx, Ly, p.w,
                                              p.P, du, ctx=ctx)
    else:
        g = pe.mex_stokes_dlp_kspace_backward(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, Lx, Ly,
                                              p.w, p.P, du, ctx=ctx)
    return {k: v.cpu().numpy() for k, v in g.items()}


KINDS = ["slp_real", "slp_kspace", "dlp_real", "dlp_kspace"]


@pytest.mark.parametrize("kind", KINDS)
def test_kept_backward_is_bitwise_the_rebuilt_one(gpu, kind):
    pe = gpu
    w = W.uniform_box(3000, seed=61, L=1.0)
    k = "slp" if kind.startswith("slp") else "dlp"
    p = Pm.select_params(k, 3000, 3000, w.f, 1.0, 1.0, tol=1e-10, M=96)
    ps, f, nv = _t(w.psrc), _t(w.f), _t(w.n)
    du = _t(np.random.default_rng(62).standard_normal((3000, 2)))
    print(f"[keep {kind}]", flush=True)
    plain = pe.Context(0)
    plain.set_adjoint_passes(2)
    keep = pe.Context(0)
    keep.set_adjoint_passes(2)
    keep.set_keep(True)
    one = pe.Context(0)
    u0 = _fwd(pe, kind, ps, f, nv, p, 1.0, 1.0, plain).cpu().numpy()
    g0 = _bwd(pe, kind, ps, f, nv, du, p, 1.0, 1.0, plain)
    u1 = _fwd(pe, kind, ps, f, nv, p, 1.0, 1.0, keep).cpu().numpy()
    g1 = _bwd(pe, kind, ps, f, nv, du, p, 1.0, 1.0, keep)
    assert np.array_equal(u0, u1)
    for name in g0:
        assert np.array_equal(g0[name], g1[name]), name
    # the default rebuild (one pass in Fourier space) agrees to rounding
    g3 = _bwd(pe, kind, ps, f, nv, du, p, 1.0, 1.0, one)
    for name in g0:
        assert relerr(g3[name], g0[name]) <= 1e-13, name
    one.close()
    # a second backward with another du reuses the same kept state
    du2 = _t(np.random.default_rng(63).standard_normal((3000, 2)))
    assert all(np.array_equal(a, b) for a, b in
               zip(_bwd(pe, kind, ps, f, nv, du2, p, 1.0, 1.0, plain).values(),
                   _bwd(pe, kind, ps, f, nv, du2, p, 1.0, 1.0, keep).values()))
    # another density array: no reuse (the kept state belongs to f)
    f2 = f * 1.5
    g2 = _bwd(pe, kind, ps, f2, nv, du, p, 1.0, 1.0, keep)
    r2 = _bwd(pe, kind, ps, f2, nv, du, p, 1.0, 1.0, plain)
    for name in r2:
        assert np.array_equal(g2[name], r2[name]), name
    keep.set_keep(False)
    keep.close()
    plain.close()


def test_kept_benchmark_adjoint_matches_oracle(gpu):
    pe = gpu
    w, p = W.dlp_ellipses_config()
    N = len(w.psrc)
    rng = np.random.default_rng(64)
    du = np.zeros((N, 2))
    sub = rng.choice(N, 2000, replace=False)
    du[sub] = rng.standard_normal((2000, 2))
    ps, f, nv, dd = _t(w.psrc), _t(w.f), _t(w.n), _t(du)
    ctx = pe.Context(0)
    ctx.set_keep(True)
    print("[keep dlp ellipses N=1e5]", flush=True)
    pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, ctx=ctx)
    pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P, ctx=ctx)
    gr = _bwd(pe, "dlp_real", ps, f, nv, dd, p, 1.0, 1.0, ctx)
    gk = _bwd(pe, "dlp_kspace", ps, f, nv, dd, p, 1.0, 1.0, ctx)
    cr = cref.dlp_real_adj(w.psrc, w.ptar, w.f, w.n, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, du)
    ck = cref.dlp_kspace_adj(w.psrc, w.ptar, p.xi, p.eta, w.f, w.n, p.Mx, p.My, 1.0, 1.0, p.w,
                             p.P, du)
    names = {"dpsrc": "psrc", "dptar": "ptar", "df": "f", "dn": "n"}
    for g, c in ((gr, cr), (gk, ck)):
        for k, v in c.items():
            e = relerr(g[names[k]], v)
            print(f"  {k}: {e:.2e}", flush=True)
            assert e <= 1e-12, k
    ctx.close()


def test_keep_mode_edge_cases(gpu):
    """Calls that keep nothing or must not reuse: separate targets and a
    replaced kept state (forward with f, forward with f2, backward with f)
    give the same results with keeping on as off."""
    pe = gpu
    w = W.uniform_box(2000, seed=65, L=1.0, targets_equal_sources=False, Nt=700)
    p = Pm.select_params("dlp", 2000, 700, w.f, 1.0, 1.0, tol=1e-10, M=64)
    ps, pt, f, nv = _t(w.psrc), _t(w.ptar), _t(w.f), _t(w.n)
    du_t = _t(np.random.default_rng(66).standard_normal((700, 2)))
    plain = pe.Context(0)
    keep = pe.Context(0)
    keep.set_keep(True)
    print("[keep edge: separate targets]", flush=True)
    outs = []
    for ctx in (plain, keep):
        ur = pe.mex_stokes_dlp_real(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, ctx=ctx)
        uk = pe.mex_stokes_dlp_kspace(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0, 1.0, p.w,
                                      p.P, ctx=ctx)
        gr = pe.mex_stokes_dlp_real_backward(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, 1.0,
                                             1.0, du_t, ctx=ctx)
        gk = pe.mex_stokes_dlp_kspace_backward(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0,
                                               1.0, p.w, p.P, du_t, ctx=ctx)
        outs.append([ur.cpu().numpy(), uk.cpu().numpy()] +
                    [v.cpu().numpy() for g in (gr, gk) for v in g.values()])
    assert all(np.array_equal(a, b) for a, b in zip(*outs))
    print("[keep edge: replaced state]", flush=True)
    q = Pm.select_params("dlp", 2000, 2000, w.f, 1.0, 1.0, tol=1e-10, M=64)
    du = _t(np.random.default_rng(67).standard_normal((2000, 2)))
    f2 = f * 0.5 + 0.25
    _fwd(pe, "dlp_real", ps, f, nv, q, 1.0, 1.0, keep)
    _fwd(pe, "dlp_kspace", ps, f, nv, q, 1.0, 1.0, keep)
    _fwd(pe, "dlp_real", ps, f2, nv, q, 1.0, 1.0, keep)     # replaces the kept state
    _fwd(pe, "dlp_kspace", ps, f2, nv, q, 1.0, 1.0, keep)
    for kind in ("dlp_real", "dlp_kspace"):
        a = _bwd(pe, kind, ps, f, nv, du, q, 1.0, 1.0, keep)
        b = _bwd(pe, kind, ps, f, nv, du, q, 1.0, 1.0, plain)
        assert all(np.array_equal(a[k], b[k]) for k in b), kind
    keep.close()
    plain.close()
