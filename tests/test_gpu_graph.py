"""Stream capture: in async mode the entry points only enqueue work on the
context's stream, so a caller can capture a whole step (real + Fourier
sums, rocFFT included) in a HIP graph and replay it.  Every replay is
checked bitwise against the eager result, with eager work and
synchronisation between replays."""
import numpy as np
import pytest
import torch

from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(gpu):
    pe = gpu
    w = W.uniform_box(2000, seed=8)
    p = Pm.select_params("dlp", 2000, 2000, w.f, 1.0, 1.0, tol=1e-10, M=64)
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ps, f, nv = t(w.psrc), t(w.f), t(w.n)
    ur = torch.empty_like(ps)
    uk = torch.empty_like(ps)
    ctx = pe.Context(0)

    def step():
        pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, ctx=ctx, out=ur)
        pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P,
                                 ctx=ctx, out=uk)

    s = torch.cuda.Stream(dev)
    with torch.cuda.stream(s):
        ctx.bind_torch_stream(s)
        ctx.set_async(True)
        step()   # allocates the workspace and the transform plans outside capture
    torch.cuda.synchronize()
    ref_r, ref_k = ur.clone(), uk.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.bind_torch_stream(torch.cuda.current_stream())
        step()
    torch.cuda.synchronize()
    for rep in range(4):
        ur.zero_()
        uk.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(ur, ref_r) and torch.equal(uk, ref_k), rep
        x = torch.randn(256, 256, device=dev) @ torch.randn(256, 256, device=dev)
        torch.cuda.synchronize()
    ctx.set_async(False)


@pytest.mark.parametrize("targets", ["mixed", "sources"])
def test_multibox_graph_replays_match_eager(gpu, targets):
    """Many boxes (real-space sums in one launch, Fourier-space sums on
    several streams forked from and joined into the captured stream; with
    targets = sources in every box, one Fourier-space sort over all boxes):
    a captured sequence replays bitwise equal to eager calls."""
    pe = gpu
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    boxes = []
    for i, (n, L) in enumerate(((700, 1.0), (3000, 2.0), (1500, 0.8), (5000, 1.5), (900, 3.0),
                                (2500, 1.2))):
        w = W.uniform_box(n, seed=40 + i, L=L)
        p = Pm.select_params("dlp", n, n, w.f, L, L, tol=1e-10)
        ps = t(w.psrc)
        same = i % 2 == 0 or targets == "sources"
        boxes.append(dict(psrc=ps, ptar=ps if same else t(w.ptar), f=t(w.f), n=t(w.n),
                          Lx=L, Ly=L, xi=p.xi, nside_x=p.nside_x, nside_y=p.nside_y, eta=p.eta,
                          w=p.w, Mx=p.Mx, My=p.My, P=p.P))
    ctx = pe.Context(0)

    def step():
        return (pe.stokes_dlp_real_multibox(boxes, ctx=ctx) +
                pe.stokes_dlp_kspace_multibox(boxes, ctx=ctx))

    s = torch.cuda.Stream(dev)
    with torch.cuda.stream(s):
        ctx.bind_torch_stream(s)
        ctx.set_async(True)
        ref = [u.clone() for u in step()]   # workspaces, slots and plans made here
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.bind_torch_stream(torch.cuda.current_stream())
        captured = step()
    torch.cuda.synchronize()
    for rep in range(4):
        for u in captured:
            u.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(u, r) for u, r in zip(captured, ref)), rep
        with torch.cuda.stream(s):
            ctx.bind_torch_stream(s)
            eager = step()   # eager work between replays
        torch.cuda.synchronize()
        assert all(torch.equal(u, r) for u, r in zip(eager, ref)), rep
    print(f"[multibox graph {targets}] {len(boxes)} boxes, 4 replays bitwise equal", flush=True)
    ctx.set_async(False)


def test_frozen_context_refuses_reallocation(gpu):
    """se2p_set_frozen: after the warm-up, the same calls still run, while a
    call that would grow the kept state, plan a new grid or free the kept
    state fails with SE2P_ERR_FROZEN."""
    from fasttools2d_amd import _capi

    pe = gpu
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    w = W.uniform_box(2000, seed=9)
    p = Pm.select_params("dlp", 2000, 2000, w.f, 1.0, 1.0, tol=1e-10, M=64)
    ps, f, nv = t(w.psrc), t(w.f), t(w.n)
    big = W.uniform_box(20000, seed=10)
    pb = Pm.select_params("dlp", 20000, 20000, big.f, 1.0, 1.0, tol=1e-10, M=128)
    bs, bf, bn = t(big.psrc), t(big.f), t(big.n)
    ctx = pe.Context(0)
    ctx.set_keep(True)
    ur = pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, ctx=ctx).clone()
    uk = pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P,
                                  ctx=ctx).clone()
    ctx.set_frozen(True)
    again = pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, ctx=ctx)
    assert torch.equal(again, ur)
    with pytest.raises(_capi.Se2pError) as e:
        pe.mex_stokes_dlp_real(bs, bs, bf, bn, pb.xi, pb.nside_x, pb.nside_y, 1.0, 1.0, ctx=ctx)
    assert e.value.code == _capi.SE2P_ERR_FROZEN
    with pytest.raises(_capi.Se2pError) as e:
        pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, 2 * p.Mx, 2 * p.My, 1.0, 1.0,
                                 p.w / 2, p.P, ctx=ctx)
    assert e.value.code == _capi.SE2P_ERR_FROZEN
    with pytest.raises(_capi.Se2pError) as e:
        ctx.set_keep(False)
    assert e.value.code == _capi.SE2P_ERR_FROZEN
    # the sized calls still run and agree with the warm-up
    again = pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P,
                                     ctx=ctx)
    assert torch.equal(again, uk)
    ctx.set_frozen(False)
    ctx.set_keep(False)
    out = pe.mex_stokes_dlp_real(bs, bs, bf, bn, pb.xi, pb.nside_x, pb.nside_y, 1.0, 1.0, ctx=ctx)
    assert torch.isfinite(out).all()


def test_stream_switch_orders_calls(gpu):
    """Calls on one context bound to two streams in turn, async: the call on
    the second stream waits for the first (both use the context's workspace),
    so both results equal synchronous calls."""
    pe = gpu
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    wa = W.uniform_box(60000, seed=11)
    pa = Pm.select_params("dlp", 60000, 60000, wa.f, 1.0, 1.0, tol=1e-10, M=256)
    wb = W.uniform_box(30000, seed=12)
    pb = Pm.select_params("dlp", 30000, 30000, wb.f, 1.0, 1.0, tol=1e-10, M=128)
    A = [t(a) for a in (wa.psrc, wa.f, wa.n)]
    B = [t(a) for a in (wb.psrc, wb.f, wb.n)]
    ref = pe.Context(0)
    ra = pe.mex_stokes_dlp_real(A[0], A[0], A[1], A[2], pa.xi, pa.nside_x, pa.nside_y, 1.0, 1.0, ctx=ref)
    rb = pe.mex_stokes_dlp_real(B[0], B[0], B[1], B[2], pb.xi, pb.nside_x, pb.nside_y, 1.0, 1.0, ctx=ref)
    torch.cuda.synchronize()
    ctx = pe.Context(0)
    s1, s2 = torch.cuda.Stream(dev), torch.cuda.Stream(dev)
    ua = torch.empty_like(ra)
    ub = torch.empty_like(rb)
    for rep in range(3):
        ua.zero_()
        ub.zero_()
        torch.cuda.synchronize()
        ctx.set_async(True)
        with torch.cuda.stream(s1):
            pe.mex_stokes_dlp_real(A[0], A[0], A[1], A[2], pa.xi, pa.nside_x, pa.nside_y, 1.0,
                                   1.0, ctx=ctx, out=ua)
        with torch.cuda.stream(s2):
            pe.mex_stokes_dlp_real(B[0], B[0], B[1], B[2], pb.xi, pb.nside_x, pb.nside_y, 1.0,
                                   1.0, ctx=ctx, out=ub)
        torch.cuda.synchronize()
        ctx.set_async(False)
        assert torch.equal(ua, ra) and torch.equal(ub, rb), rep
