"""Concurrent groups (se2p_group_begin / se2p_group_end): calls made between
them run side by side on stream slots of their own and give bitwise the
results of the same calls made one after the other -- forward sums of both
kernels (targets = sources and distinct targets), batched and backward
calls, a kept forward state made in a group and used by backward calls after
it, graph capture of a grouped step (every replay checked against the eager
result) -- and the calls and settings a group does not allow fail with
SE2P_ERR_ARG without disturbing the context."""
import numpy as np
import pytest
import torch

from fasttools2d_amd import _capi
from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _setup(n, seed, kind="dlp", same=True, L=1.0):
    w = W.uniform_box(n, seed=seed, L=L)
    p = Pm.select_params(kind, n, n, w.f, L, L, tol=1e-10)
    ps = _t(w.psrc)
    pt = ps if same else _t(w.ptar)
    return w, p, ps, pt, _t(w.f), _t(w.n)


def _calls(pe, ctx, p, ps, pt, f, nv, L):
    """The four forward sums on one geometry, each a closure returning its
    output tensor."""
    return [
        lambda: pe.mex_stokes_slp_kspace(ps, pt, p.xi, p.eta, f, p.Mx, p.My, L, L, p.w, p.P,
                                         ctx=ctx),
        lambda: pe.mex_stokes_slp_real(ps, pt, f, p.xi, p.nside_x, p.nside_y, L, L, ctx=ctx),
        lambda: pe.mex_stokes_dlp_kspace(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, L, L, p.w, p.P,
                                         ctx=ctx),
        lambda: pe.mex_stokes_dlp_real(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, L, L, ctx=ctx),
    ]


@pytest.mark.parametrize("same", [True, False])
def test_grouped_forward_calls_are_bitwise_sequential(gpu, same):
    pe = gpu
    L = 1.3
    w, p, ps, pt, f, nv = _setup(6000, 91, same=same, L=L)
    ctx = pe.Context(0)
    calls = _calls(pe, ctx, p, ps, pt, f, nv, L)
    ref = [c().clone() for c in calls]
    for rep in range(3):   # the first group creates the slots, later ones reuse them
        with ctx.group():
            got = [c() for c in calls]
        for i, (a, b) in enumerate(zip(got, ref)):
            assert torch.equal(a, b), (rep, i)
    # async mode: group_end only joins; the context stream orders the reads
    ctx.set_async(True)
    with ctx.group():
        got = [c() for c in calls[:2]]
    torch.cuda.synchronize()
    ctx.set_async(False)
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])
    ctx.close()


def test_grouped_benchmark_pair_is_bitwise_sequential(gpu):
    """The benchmark's step (stresslet, 200 ellipses x 500 nodes, grid 512^2):
    the Fourier-space call on the context stream and the real-space call on a
    slot, side by side."""
    pe = gpu
    w, p = W.dlp_ellipses_config(seed=2, n_ell=200, n_nodes=500, tol=1e-10, M=512)
    ps, f, nv = _t(w.psrc), _t(w.f), _t(w.n)
    ctx = pe.Context(0)

    def real():
        return pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, w.Lx, w.Ly,
                                      ctx=ctx)

    def fourier():
        return pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, w.Lx, w.Ly, p.w,
                                        p.P, ctx=ctx)

    ur, uk = real(), fourier()
    for order in ("fourier first", "real first"):
        with ctx.group():
            if order == "fourier first":
                gk, gr = fourier(), real()
            else:
                gr, gk = real(), fourier()
        assert torch.equal(gr, ur) and torch.equal(gk, uk), order
    ctx.close()


def test_grouped_batched_and_backward_are_bitwise_sequential(gpu):
    pe = gpu
    L = 1.0
    w, p, ps, pt, f, nv = _setup(4000, 92, kind="slp")
    F = _t(np.random.default_rng(3).standard_normal((5, len(w.psrc), 2)))
    du = _t(np.random.default_rng(4).standard_normal((len(w.psrc), 2)))
    ctx = pe.Context(0)
    calls = [
        lambda: pe.stokes_slp_kspace_batched(ps, ps, p.xi, p.eta, F, p.Mx, p.My, L, L, p.w, p.P,
                                             ctx=ctx),
        lambda: pe.stokes_slp_real_batched(ps, ps, F, p.xi, p.nside_x, p.nside_y, L, L, ctx=ctx),
        lambda: pe.mex_stokes_slp_kspace_backward(ps, ps, p.xi, p.eta, f, p.Mx, p.My, L, L, p.w,
                                                  p.P, du, ctx=ctx),
        lambda: pe.mex_stokes_slp_real_backward(ps, ps, f, p.xi, p.nside_x, p.nside_y, L, L, du,
                                                ctx=ctx),
    ]

    def snap(x):
        return {k: v.clone() for k, v in x.items()} if isinstance(x, dict) else x.clone()

    ref = [snap(c()) for c in calls]
    with ctx.group():
        got = [c() for c in calls]
    for i, (a, b) in enumerate(zip(got, ref)):
        if isinstance(b, dict):
            assert a.keys() == b.keys()
            for k in b:
                assert torch.equal(a[k], b[k]), (i, k)
        else:
            assert torch.equal(a, b), i
    ctx.close()


def test_kept_state_made_in_a_group_serves_backward_after_it(gpu):
    pe = gpu
    L = 1.0
    w, p, ps, pt, f, nv = _setup(5000, 93)
    du = _t(np.random.default_rng(5).standard_normal((len(w.psrc), 2)))
    ref_ctx = pe.Context(0)
    ref_ctx.set_keep(True)
    ref_fwd = [pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, L, L, p.w, p.P,
                                        ctx=ref_ctx),
               pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, L, L,
                                      ctx=ref_ctx)]
    ref_bwd = [pe.mex_stokes_dlp_kspace_backward(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, L, L,
                                                 p.w, p.P, du, ctx=ref_ctx),
               pe.mex_stokes_dlp_real_backward(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, L, L,
                                               du, ctx=ref_ctx)]
    ref_ctx.close()
    ctx = pe.Context(0)
    ctx.set_keep(True)
    with ctx.group():
        fwd = [pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, L, L, p.w, p.P,
                                        ctx=ctx),
               pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, L, L, ctx=ctx)]
    for rep in range(2):   # backward calls grouped, then not
        if rep == 0:
            with ctx.group():
                bwd = [pe.mex_stokes_dlp_kspace_backward(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My,
                                                         L, L, p.w, p.P, du, ctx=ctx),
                       pe.mex_stokes_dlp_real_backward(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y,
                                                       L, L, du, ctx=ctx)]
        else:
            bwd = [pe.mex_stokes_dlp_kspace_backward(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, L, L,
                                                     p.w, p.P, du, ctx=ctx),
                   pe.mex_stokes_dlp_real_backward(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, L,
                                                   L, du, ctx=ctx)]
        for a, b in zip(fwd, ref_fwd):
            assert torch.equal(a, b)
        for a, b in zip(bwd, ref_bwd):
            for k in b:
                assert torch.equal(a[k], b[k]), (rep, k)
    ctx.set_keep(False)
    ctx.close()


def test_graph_of_a_grouped_step_replays_match_eager(gpu):
    pe = gpu
    w, p, ps, pt, f, nv = _setup(3000, 94)
    dev = torch.device("cuda", 0)
    ur = torch.empty_like(ps)
    uk = torch.empty_like(ps)
    ctx = pe.Context(0)

    def step():
        with ctx.group():
            pe.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P,
                                     ctx=ctx, out=uk)
            pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0, ctx=ctx,
                                   out=ur)

    s = torch.cuda.Stream(dev)
    with torch.cuda.stream(s):
        ctx.bind_torch_stream(s)
        ctx.set_async(True)
        step()   # sizes the workspaces, slots and plans outside capture
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
    del g
    ctx.close()


def test_group_refusals_leave_the_context_usable(gpu):
    pe = gpu
    L = 1.0
    w, p, ps, pt, f, nv = _setup(2000, 95)
    ctx = pe.Context(0)
    ref = pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, L, L, ctx=ctx)
    lib, h = ctx._lib, ctx.handle
    assert lib.se2p_group_end(h) == _capi.SE2P_ERR_ARG          # none open
    assert lib.se2p_group_begin(h) == 0
    assert lib.se2p_group_begin(h) == _capi.SE2P_ERR_ARG        # nested
    assert lib.se2p_set_keep(h, 0) == _capi.SE2P_ERR_ARG
    other = torch.cuda.Stream()
    assert lib.se2p_set_stream(h, _capi.ctypes.c_void_p(other.cuda_stream)) == _capi.SE2P_ERR_ARG
    assert lib.se2p_set_spread_chunk(h, 128) == _capi.SE2P_ERR_ARG
    with pytest.raises(_capi.Se2pError):
        pe.count_pairs(ps, ps, p.nside_x, p.nside_y, L, L, ctx=ctx)
    box = dict(psrc=ps, ptar=ps, f=f, n=nv, Lx=L, Ly=L, xi=p.xi, nside_x=p.nside_x,
               nside_y=p.nside_y, eta=p.eta, w=p.w, Mx=p.Mx, My=p.My, P=p.P)
    with pytest.raises(_capi.Se2pError):
        pe.stokes_dlp_real_multibox([box], ctx=ctx)
    ctx.set_host_pointers(True)
    hp, hf, hn = (np.ascontiguousarray(a, dtype=np.float64) for a in (w.psrc, w.f, w.n))
    u_host = np.empty((len(w.psrc), 2))
    vp = lambda a: a.ctypes.data_as(_capi.ctypes.c_void_p)
    rc = lib.se2p_dlp_real(h, vp(hp), len(hp), vp(hp), len(hp), vp(hf), vp(hn), p.xi,
                           p.nside_x, p.nside_y, L, L, vp(u_host))
    assert rc == _capi.SE2P_ERR_ARG
    ctx.set_host_pointers(False)
    outs = [pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, L, L, ctx=ctx)
            for _ in range(4)]
    with pytest.raises(_capi.Se2pError):   # a fifth call
        pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, L, L, ctx=ctx)
    assert lib.se2p_group_end(h) == 0
    for u in outs:
        assert torch.equal(u, ref)
    # outside the group everything works again
    assert torch.equal(pe.mex_stokes_dlp_real(ps, ps, f, nv, p.xi, p.nside_x, p.nside_y, L, L,
                                              ctx=ctx), ref)
    pe.count_pairs(ps, ps, p.nside_x, p.nside_y, L, L, ctx=ctx)
    ctx.close()
