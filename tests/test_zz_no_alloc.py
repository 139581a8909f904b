"""se2p_dlp_real_noalloc through the C ABI: the stresslet real-space sum with
its sorted state in the library's fixed device scratch instead of allocated
workspace.  Checked bitwise against se2p_dlp_real (the same kernels) on two
calls in a row, the second smaller than the first (the scratch keeps the
first call's contents), against the oracle at 1e-12, and that the calls leave
the device's free memory unchanged."""
import ctypes

import numpy as np
import pytest
import torch

from fasttools2d_amd import _capi
from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W
from oracle import cref

pytestmark = pytest.mark.gpu

D = ctypes.c_void_p


def _d(t):
    return D(t.data_ptr() if t.numel() else 0)


def _case(N, seed, nt=None, M=64):
    w = W.uniform_box(N, seed=seed, targets_equal_sources=nt is None, Nt=nt)
    p = Pm.select_params("dlp", N, nt or N, w.f, 1.0, 1.0, tol=1e-10, M=M)
    return w, p


def _run(lib, h, entry, dev, w, p, same):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ps, f, nv = t(w.psrc), t(w.f), t(w.n)
    pt = ps if same else t(w.ptar)
    u = torch.full((pt.shape[0], 2), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    rc = getattr(lib, entry)(h, _d(ps), ps.shape[0], _d(pt), pt.shape[0], _d(f), _d(nv), p.xi,
                             p.nside_x, p.nside_y, 1.0, 1.0, _d(u))
    assert rc == 0, (entry, lib.se2p_last_error(h))
    return u.cpu().numpy()


def test_noalloc_matches_plain_entry_two_calls(gpu):
    dev = torch.device("cuda", 0)
    lib = _capi.load()
    h = ctypes.c_void_p()
    assert lib.se2p_create(ctypes.byref(h), 0) == 0
    try:
        cases = [
            ("large", *_case(60000, 31, M=256), True),     # targets are the sources
            ("smaller", *_case(3000, 32), True),           # second call, smaller input
            ("separate targets", *_case(2500, 33, nt=1700), False),
        ]
        for name, w, p, same in cases:
            print(f"noalloc case {name}: Ns={w.psrc.shape[0]} nside={p.nside_x}", flush=True)
            u_na = _run(lib, h, "se2p_dlp_real_noalloc", dev, w, p, same)
            u_ref = _run(lib, h, "se2p_dlp_real", dev, w, p, same)
            assert np.array_equal(u_na, u_ref), name
            if w.psrc.shape[0] <= 3000:
                pt = w.psrc if same else w.ptar
                c = cref.dlp_real(w.psrc, pt, w.f, w.n, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
                assert np.abs(u_na - c).max() <= 1e-12 * np.abs(c).max(), name
    finally:
        lib.se2p_destroy(h)


def test_noalloc_allocates_nothing(gpu):
    """Free device memory is unchanged across no-allocation calls on a fresh
    context, while the plain entry on the same fresh state allocates its
    workspace (so the probe sees an allocation of this size)."""
    dev = torch.device("cuda", 0)
    lib = _capi.load()
    w1, p1 = _case(40000, 41, M=128)
    w2, p2 = _case(5000, 42)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    arrays = []
    for w in (w1, w2):
        ps = t(w.psrc)
        arrays.append((ps, t(w.f), t(w.n),
                       torch.empty((ps.shape[0], 2), dtype=torch.float64, device=dev)))
    h = ctypes.c_void_p()
    assert lib.se2p_create(ctypes.byref(h), 0) == 0
    try:
        def call(entry, k, p):
            ps, f, nv, u = arrays[k]
            rc = getattr(lib, entry)(h, _d(ps), ps.shape[0], _d(ps), ps.shape[0], _d(f), _d(nv),
                                     p.xi, p.nside_x, p.nside_y, 1.0, 1.0, _d(u))
            assert rc == 0, lib.se2p_last_error(h)

        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info(0)[0]
        call("se2p_dlp_real_noalloc", 0, p1)
        call("se2p_dlp_real_noalloc", 1, p2)
        call("se2p_dlp_real_noalloc", 0, p1)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info(0)[0]
        assert free1 == free0, (free0, free1)
        # the plain entry on this context grows its workspace
        call("se2p_dlp_real", 0, p1)
        torch.cuda.synchronize()
        free2 = torch.cuda.mem_get_info(0)[0]
        assert free2 < free1, (free1, free2)
    finally:
        lib.se2p_destroy(h)


def test_noalloc_rejects_oversize_and_host_pointers(gpu):
    lib = _capi.load()
    h = ctypes.c_void_p()
    assert lib.se2p_create(ctypes.byref(h), 0) == 0
    try:
        fake = D(256)   # never dereferenced: the size check comes first
        rc = lib.se2p_dlp_real_noalloc(h, fake, 400000, fake, 400000, fake, fake, 100.0, 22, 22,
                                       1.0, 1.0, fake)
        assert rc == _capi.SE2P_ERR_ARG
        assert b"fixed scratch holds" in lib.se2p_last_error(h)
        assert lib.se2p_set_host_pointers(h, 1) == 0
        x = np.zeros((10, 2))
        p = D(x.ctypes.data)
        rc = lib.se2p_dlp_real_noalloc(h, p, 10, p, 10, p, p, 10.0, 1, 1, 1.0, 1.0, p)
        assert rc == _capi.SE2P_ERR_ARG
        assert b"device pointers only" in lib.se2p_last_error(h)
    finally:
        lib.se2p_destroy(h)
