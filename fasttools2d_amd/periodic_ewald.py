"""Host-side mirror of the reference's PeriodicEwald interface.

The four ``mex_stokes_*`` functions take the same arguments, in the same
order, as the reference MEX kernels and call the HIP implementation through
the C ABI (``include/se2p.h``):

* ``mex_stokes_slp_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly)``
  -- PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_real.cpp:58-83
* ``mex_stokes_slp_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P)``
  -- PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_kspace.cpp:8-42
* ``mex_stokes_dlp_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly)``
  -- PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_real.cpp:15-52
* ``mex_stokes_dlp_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P)``
  -- PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_kspace.cpp:10-53

Arrays: a MATLAB ``2 x N`` matrix is an ``(N, 2)`` float64 array here (the
same interleaved memory).  Device tensors (``torch`` on ``cuda``) are used in
place; numpy arrays or CPU tensors are copied to the current device and the
result is returned as a device tensor.  Errors raise ``ValueError`` with the
reference's message for the checks the MEX performs, and ``Se2pError`` for
codes returned by the C ABI.
"""
from __future__ import annotations

import contextlib
import ctypes
from typing import Optional

import numpy as np
import torch

from . import _capi


class Context:
    """Owns an ``se2p_ctx`` (stream, workspace, FFT plans) on one device."""

    def __init__(self, device: Optional[int] = None):
        if device is None:
            device = torch.cuda.current_device()
        self.device = int(device)
        self._lib = _capi.load()
        h = ctypes.c_void_p()
        rc = self._lib.se2p_create(ctypes.byref(h), self.device)
        if rc != 0:
            raise _capi.Se2pError(rc, f"se2p_create failed on device {self.device}")
        self.handle = h
        self._stream_bound = None

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.se2p_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stream / timing -----------------------------------------------------
    def bind_torch_stream(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        """Run on torch's stream so inputs written by torch are ordered before
        the library's kernels and outputs are visible to torch afterwards."""
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        ptr = stream.cuda_stream
        if ptr == 0:
            # torch is on the legacy null stream: make its pending writes
            # visible, then keep the context's own stream
            stream.synchronize()
        if ptr != self._stream_bound:
            self._check(self._lib.se2p_set_stream(self.handle, ctypes.c_void_p(ptr)))
            self._stream_bound = ptr

    def set_async(self, flag: bool) -> None:
        self._check(self._lib.se2p_set_async(self.handle, 1 if flag else 0))

    def set_frozen(self, flag: bool) -> None:
        """Refuse (SE2P_ERR_FROZEN) every call that would allocate or free
        device memory, e.g. while HIP graphs captured from this context
        exist (se2p_set_frozen)."""
        self._check(self._lib.se2p_set_frozen(self.handle, 1 if flag else 0))

    def set_timing(self, flag: bool) -> None:
        self._check(self._lib.se2p_set_timing(self.handle, 1 if flag else 0))

    def set_host_pointers(self, flag: bool) -> None:
        """Array arguments of raw C-ABI calls on this context are host memory
        (the library stages them).  The Python wrappers of this module pass
        device tensors and need the default (False)."""
        self._check(self._lib.se2p_set_host_pointers(self.handle, 1 if flag else 0))

    SORT_RADIX = 1       # force the multi-pass radix sort
    SORT_BOX_KERNEL = 2  # separate run-bounding-box kernel

    def set_keep(self, flag: bool) -> None:
        """Keep the sorted state of single-density, targets-are-sources
        forward calls for the matching backward call (se2p_set_keep).  The
        forward's arrays must stay unchanged until the backward call."""
        self._check(self._lib.se2p_set_keep(self.handle, int(bool(flag))))

    def set_spread_chunk(self, chunk: int) -> None:
        """Candidate sources per spreading workgroup (se2p_set_spread_chunk):
        0 for the default, else a multiple of 16 in [64, 2048].  Results for
        different values differ by rounding."""
        self._check(self._lib.se2p_set_spread_chunk(self.handle, int(chunk)))

    def set_sort_mode(self, mode: int) -> None:
        """Sorting of cell lists and bins: 0 (default) uses the fused sort
        where it fits; SORT_RADIX / SORT_BOX_KERNEL force the separate
        kernels.  Every mode gives bitwise identical results."""
        self._check(self._lib.se2p_set_sort_mode(self.handle, int(mode)))

    def set_adjoint_passes(self, passes: int) -> None:
        """Fourier-space backward calls with targets = sources and both
        sides wanted: 1 (default) runs both passes as one, 2 one after the
        other (se2p_set_adjoint_passes).  Results agree to rounding."""
        self._check(self._lib.se2p_set_adjoint_passes(self.handle, int(passes)))

    def group_begin(self) -> None:
        """Open a concurrent group (se2p_group_begin): each following call
        runs on a stream slot of its own, side by side with the others,
        until group_end.  The calls must be independent.  The context is
        bound to torch's current stream first (a group cannot change it)."""
        self.bind_torch_stream(torch.cuda.current_stream(self.device))
        self._check(self._lib.se2p_group_begin(self.handle))
        self._hold = []
        _group_holds.append(self._hold)

    def group_end(self) -> None:
        """Join the group's calls into the context stream (se2p_group_end);
        their outputs are defined from here on."""
        hold = getattr(self, "_hold", None)
        try:
            self._check(self._lib.se2p_group_end(self.handle))
        finally:
            if hold is not None:
                # by identity: == on lists would compare the tensors
                _group_holds[:] = [h for h in _group_holds if h is not hold]
                self._hold = None

    @contextlib.contextmanager
    def group(self):
        """``with ctx.group(): ...`` -- group_begin, the calls, group_end
        (also when a call raises, so that the context stays usable)."""
        self.group_begin()
        try:
            yield self
        finally:
            self.group_end()

    def stage_times(self) -> dict:
        buf = (ctypes.c_float * len(_capi.STAGES))()
        self._check(self._lib.se2p_stage_times(self.handle, buf, len(_capi.STAGES)))
        return {name: float(buf[i]) for i, name in enumerate(_capi.STAGES)}

    def last_error(self) -> str:
        return (self._lib.se2p_last_error(self.handle) or b"").decode()

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise _capi.Se2pError(rc, self.last_error())


_contexts: dict = {}


def get_context(device: Optional[int] = None) -> Context:
    if device is None:
        device = torch.cuda.current_device()
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        _contexts[device] = ctx
    return ctx


# -- argument handling ---------------------------------------------------------
def _as_points(a, name: str, device) -> torch.Tensor:
    """(N, 2) float64 contiguous device tensor; MATLAB's 2xN check."""
    if isinstance(a, np.ndarray):
        a = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))
    if not isinstance(a, torch.Tensor):
        a = torch.as_tensor(a, dtype=torch.float64)
    if a.dim() != 2 or a.shape[1] != 2:
        raise ValueError(f"{name} must be a 2xn matrix.")
    a = a.to(device=device, dtype=torch.float64).contiguous()
    return a


def _output(out, nt: int, device) -> torch.Tensor:
    """The result tensor: a new (Nt, 2) float64 device tensor, or the caller's
    ``out``, which the library overwrites through its data pointer and which
    must therefore be exactly that: contiguous float64 (Nt, 2) on the device
    of the inputs."""
    if out is None:
        return torch.empty((nt, 2), dtype=torch.float64, device=device)
    if not isinstance(out, torch.Tensor):
        raise ValueError("out must be a torch tensor.")
    if out.dtype != torch.float64 or tuple(out.shape) != (nt, 2):
        raise ValueError(f"out must be a float64 tensor of shape ({nt}, 2).")
    if out.device != torch.device(device):
        raise ValueError(f"out must be on {device}.")
    if not out.is_contiguous():
        raise ValueError("out must be contiguous.")
    return out


def _same_size(ps, a, name: str) -> None:
    """The reference's size check of a per-source array (any array type:
    checked before anything is moved to the device)."""
    if a is not None and np.shape(a)[:1] != np.shape(ps)[:1]:
        raise ValueError(f"psrc and {name} must be the same size.")


# Tensors handed to the library while a group is open: a grouped call reads
# and writes them on a stream slot after the wrapper has returned, so they
# (temporary copies of numpy inputs included) stay referenced until
# group_end; otherwise torch's allocator could hand their memory to the next
# call of the group while the first still uses it.
_group_holds: list = []


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    for hold in _group_holds:
        hold.append(t)
    return ctypes.c_void_p(t.data_ptr() if t.numel() > 0 else 0)


def _device_of(*arrays) -> torch.device:
    for a in arrays:
        if isinstance(a, torch.Tensor) and a.is_cuda:
            return a.device
    return torch.device("cuda", torch.cuda.current_device())


def _prepare(ctx: Optional[Context], dev: torch.device) -> Context:
    if ctx is None:
        ctx = get_context(dev.index)
    ctx.bind_torch_stream(torch.cuda.current_stream(dev))
    return ctx


# -- the four MEX equivalents -------------------------------------------------
def mex_stokes_slp_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, *, ctx=None, out=None):
    dev = _device_of(psrc, ptar, f)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    ff = _as_points(f, "f", dev)
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    u = _output(out, pt.shape[0], dev)
    ctx = _prepare(ctx, dev)
    ctx._check(ctx._lib.se2p_slp_real(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0],
                                      _ptr(ff), float(xi), int(nside_x), int(nside_y),
                                      float(Lx), float(Ly), _ptr(u)))
    return u


def mex_stokes_slp_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P, *, ctx=None, out=None):
    dev = _device_of(psrc, ptar, f)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    ff = _as_points(f, "f", dev)
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    u = _output(out, pt.shape[0], dev)
    ctx = _prepare(ctx, dev)
    ctx._check(ctx._lib.se2p_slp_kspace(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0],
                                        float(xi), float(eta), _ptr(ff), int(Mx), int(My),
                                        float(Lx), float(Ly), float(w), int(P), _ptr(u)))
    return u


def mex_stokes_dlp_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, *, ctx=None, out=None):
    dev = _device_of(psrc, ptar, f, n)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    ff, nn = _as_points(f, "f", dev), _as_points(n, "n", dev)
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    if nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    u = _output(out, pt.shape[0], dev)
    ctx = _prepare(ctx, dev)
    ctx._check(ctx._lib.se2p_dlp_real(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0],
                                      _ptr(ff), _ptr(nn), float(xi), int(nside_x), int(nside_y),
                                      float(Lx), float(Ly), _ptr(u)))
    return u


def mex_stokes_dlp_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P, *, ctx=None,
                          out=None):
    dev = _device_of(psrc, ptar, f, n)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    ff, nn = _as_points(f, "f", dev), _as_points(n, "n", dev)
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    if nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    u = _output(out, pt.shape[0], dev)
    ctx = _prepare(ctx, dev)
    ctx._check(ctx._lib.se2p_dlp_kspace(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0],
                                        float(xi), float(eta), _ptr(ff), _ptr(nn), int(Mx), int(My),
                                        float(Lx), float(Ly), float(w), int(P), _ptr(u)))
    return u


def count_pairs(psrc, ptar, nside_x, nside_y, Lx, Ly, *, ctx=None):
    """(pairs_in, pairs_out) of the real-space sum for this geometry."""
    dev = _device_of(psrc, ptar)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    ctx = _prepare(ctx, dev)
    a, b = ctypes.c_int64(), ctypes.c_int64()
    ctx._check(ctx._lib.se2p_count_pairs(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0],
                                         int(nside_x), int(nside_y), float(Lx), float(Ly),
                                         ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


# -- batched right-hand sides (one geometry, many densities) -----------------
def _as_batch(F, name, ns, device):
    """(nrhs, N, 2) float64 contiguous device tensor."""
    if isinstance(F, np.ndarray):
        F = torch.from_numpy(np.ascontiguousarray(F, dtype=np.float64))
    if F.dim() != 3 or F.shape[2] != 2 or F.shape[1] != ns:
        raise ValueError(f"{name} must be (nrhs, N, 2) with N = number of sources.")
    return F.to(device=device, dtype=torch.float64).contiguous()


def stokes_slp_real_batched(psrc, ptar, F, xi, nside_x, nside_y, Lx, Ly, *, ctx=None):
    """mex_stokes_slp_real for every density F[r]; returns (nrhs, Nt, 2)."""
    dev = _device_of(psrc, ptar, F)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    FF = _as_batch(F, "F", ps.shape[0], dev)
    ctx = _prepare(ctx, dev)
    nr = FF.shape[0]
    u = torch.empty((nr, pt.shape[0], 2), dtype=torch.float64, device=dev)
    ctx._check(ctx._lib.se2p_slp_real_batched(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt),
                                              pt.shape[0], _ptr(FF), nr, 2 * ps.shape[0],
                                              float(xi), int(nside_x), int(nside_y), float(Lx),
                                              float(Ly), _ptr(u), 2 * pt.shape[0]))
    return u


def stokes_slp_kspace_batched(psrc, ptar, xi, eta, F, Mx, My, Lx, Ly, w, P, *, ctx=None):
    """mex_stokes_slp_kspace for every density F[r]; returns (nrhs, Nt, 2)."""
    dev = _device_of(psrc, ptar, F)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    FF = _as_batch(F, "F", ps.shape[0], dev)
    ctx = _prepare(ctx, dev)
    nr = FF.shape[0]
    u = torch.empty((nr, pt.shape[0], 2), dtype=torch.float64, device=dev)
    ctx._check(ctx._lib.se2p_slp_kspace_batched(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt),
                                                pt.shape[0], float(xi), float(eta), _ptr(FF), nr,
                                                2 * ps.shape[0], int(Mx), int(My), float(Lx),
                                                float(Ly), float(w), int(P), _ptr(u),
                                                2 * pt.shape[0]))
    return u


def stokes_dlp_real_batched(psrc, ptar, F, n, xi, nside_x, nside_y, Lx, Ly, *, ctx=None):
    """mex_stokes_dlp_real for every density F[r] (normals shared)."""
    dev = _device_of(psrc, ptar, F, n)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    nn = _as_points(n, "n", dev)
    FF = _as_batch(F, "F", ps.shape[0], dev)
    if nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    ctx = _prepare(ctx, dev)
    nr = FF.shape[0]
    u = torch.empty((nr, pt.shape[0], 2), dtype=torch.float64, device=dev)
    ctx._check(ctx._lib.se2p_dlp_real_batched(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt),
                                              pt.shape[0], _ptr(FF), _ptr(nn), nr,
                                              2 * ps.shape[0], float(xi), int(nside_x),
                                              int(nside_y), float(Lx), float(Ly), _ptr(u),
                                              2 * pt.shape[0]))
    return u


def stokes_dlp_kspace_batched(psrc, ptar, xi, eta, F, n, Mx, My, Lx, Ly, w, P, *, ctx=None):
    """mex_stokes_dlp_kspace for every density F[r] (normals shared)."""
    dev = _device_of(psrc, ptar, F, n)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    nn = _as_points(n, "n", dev)
    FF = _as_batch(F, "F", ps.shape[0], dev)
    if nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    ctx = _prepare(ctx, dev)
    nr = FF.shape[0]
    u = torch.empty((nr, pt.shape[0], 2), dtype=torch.float64, device=dev)
    ctx._check(ctx._lib.se2p_dlp_kspace_batched(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt),
                                                pt.shape[0], float(xi), float(eta), _ptr(FF),
                                                _ptr(nn), nr, 2 * ps.shape[0], int(Mx), int(My),
                                                float(Lx), float(Ly), float(w), int(P), _ptr(u),
                                                2 * pt.shape[0]))
    return u


# -- many independent boxes in one call -------------------------------------------
_BOX_KEYS = ("psrc", "ptar", "f", "n", "Lx", "Ly", "xi", "nside_x", "nside_y", "eta", "w", "Mx",
             "My", "P")


def _multibox(name, boxes, dlp, ctx):
    """boxes: sequence of dicts with keys among psrc, ptar, f, n (stresslet),
    Lx, Ly, xi, nside_x, nside_y (real space), eta, w, Mx, My, P (Fourier
    space).  Returns the list of (Nt, 2) results, one per box."""
    arrays = [b[k] for b in boxes for k in ("psrc", "ptar", "f", "n") if b.get(k) is not None]
    dev = _device_of(*arrays)
    keep, outs = [], []
    descs = (_capi.Se2pBoxDesc * max(1, len(boxes)))()
    for i, b in enumerate(boxes):
        unknown = set(b) - set(_BOX_KEYS)
        if unknown:
            raise ValueError(f"box {i}: unknown keys {sorted(unknown)}")
        ps, pt = _as_points(b["psrc"], "psrc", dev), _as_points(b["ptar"], "ptar", dev)
        ff = _as_points(b["f"], "f", dev)
        if ff.shape[0] != ps.shape[0]:
            raise ValueError("psrc and f must be the same size.")
        nn = None
        if dlp:
            nn = _as_points(b["n"], "n", dev)
            if nn.shape[0] != ps.shape[0]:
                raise ValueError("psrc and n must be the same size.")
        u = torch.empty((pt.shape[0], 2), dtype=torch.float64, device=dev)
        keep += [ps, pt, ff, nn]
        outs.append(u)
        d = descs[i]
        d.psrc, d.ns = _ptr(ps), ps.shape[0]
        d.ptar, d.nt = _ptr(pt), pt.shape[0]
        d.f = _ptr(ff)
        d.n = _ptr(nn) if nn is not None else None
        d.Lx, d.Ly = float(b["Lx"]), float(b["Ly"])
        d.xi = float(b["xi"])
        d.nside_x, d.nside_y = int(b.get("nside_x", 1)), int(b.get("nside_y", 1))
        d.eta, d.w = float(b.get("eta", 0.0)), float(b.get("w", 0.0))
        d.Mx, d.My, d.P = int(b.get("Mx", 0)), int(b.get("My", 0)), int(b.get("P", 0))
        d.u = _ptr(u)
    ctx = _prepare(ctx, dev)
    ctx._check(getattr(ctx._lib, name)(ctx.handle, descs, len(boxes)))
    return outs


def stokes_slp_real_multibox(boxes, *, ctx=None):
    """mex_stokes_slp_real on every box (list of box dicts), one fused launch."""
    return _multibox("se2p_slp_real_multibox", boxes, False, ctx)


def stokes_slp_kspace_multibox(boxes, *, ctx=None):
    """mex_stokes_slp_kspace on every box."""
    return _multibox("se2p_slp_kspace_multibox", boxes, False, ctx)


def stokes_dlp_real_multibox(boxes, *, ctx=None):
    """mex_stokes_dlp_real on every box, one fused launch."""
    return _multibox("se2p_dlp_real_multibox", boxes, True, ctx)


def stokes_dlp_kspace_multibox(boxes, *, ctx=None):
    """mex_stokes_dlp_kspace on every box."""
    return _multibox("se2p_dlp_kspace_multibox", boxes, True, ctx)


# -- adjoints (backward entry points) --------------------------------------------
_WRT_ALL = ("psrc", "ptar", "f", "n")


def _grad_outputs(ns, nt, dev, wrt, dlp):
    out = {}
    for name, rows in (("psrc", ns), ("ptar", nt), ("f", ns), ("n", ns)):
        if name == "n" and not dlp:
            continue
        if name in wrt:
            out[name] = torch.empty((rows, 2), dtype=torch.float64, device=dev)
    return out


def _optr(out, name):
    t = out.get(name)
    return _ptr(t) if t is not None else ctypes.c_void_p(0)


def _adjoint_inputs(psrc, ptar, f, n, du, dlp):
    dev = _device_of(psrc, ptar, f, n, du)
    ps, pt = _as_points(psrc, "psrc", dev), _as_points(ptar, "ptar", dev)
    ff = _as_points(f, "f", dev)
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    nn = None
    if dlp:
        nn = _as_points(n, "n", dev)
        if nn.shape[0] != ps.shape[0]:
            raise ValueError("psrc and n must be the same size.")
    dd = _as_points(du, "du", dev)
    if dd.shape[0] != pt.shape[0]:
        raise ValueError("ptar and du must be the same size.")
    return dev, ps, pt, ff, nn, dd


def mex_stokes_slp_real_backward(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, du, *,
                                 wrt=_WRT_ALL, ctx=None):
    """Gradients of sum(du * mex_stokes_slp_real(...)): dict with keys among
    'psrc', 'ptar', 'f' (those named in wrt)."""
    dev, ps, pt, ff, _, dd = _adjoint_inputs(psrc, ptar, f, None, du, False)
    ctx = _prepare(ctx, dev)
    g = _grad_outputs(ps.shape[0], pt.shape[0], dev, wrt, False)
    ctx._check(ctx._lib.se2p_slp_real_backward(
        ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0], _ptr(ff), float(xi),
        int(nside_x), int(nside_y), float(Lx), float(Ly), _ptr(dd), _optr(g, "psrc"),
        _optr(g, "ptar"), _optr(g, "f")))
    return g


def mex_stokes_slp_kspace_backward(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P, du, *,
                                   wrt=_WRT_ALL, ctx=None):
    """Gradients of sum(du * mex_stokes_slp_kspace(...)) (dict, as above)."""
    dev, ps, pt, ff, _, dd = _adjoint_inputs(psrc, ptar, f, None, du, False)
    ctx = _prepare(ctx, dev)
    g = _grad_outputs(ps.shape[0], pt.shape[0], dev, wrt, False)
    ctx._check(ctx._lib.se2p_slp_kspace_backward(
        ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0], float(xi), float(eta),
        _ptr(ff), int(Mx), int(My), float(Lx), float(Ly), float(w), int(P), _ptr(dd),
        _optr(g, "psrc"), _optr(g, "ptar"), _optr(g, "f")))
    return g


def mex_stokes_dlp_real_backward(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, du, *,
                                 wrt=_WRT_ALL, ctx=None):
    """Gradients of sum(du * mex_stokes_dlp_real(...)): keys among 'psrc',
    'ptar', 'f', 'n'."""
    dev, ps, pt, ff, nn, dd = _adjoint_inputs(psrc, ptar, f, n, du, True)
    ctx = _prepare(ctx, dev)
    g = _grad_outputs(ps.shape[0], pt.shape[0], dev, wrt, True)
    ctx._check(ctx._lib.se2p_dlp_real_backward(
        ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0], _ptr(ff), _ptr(nn), float(xi),
        int(nside_x), int(nside_y), float(Lx), float(Ly), _ptr(dd), _optr(g, "psrc"),
        _optr(g, "ptar"), _optr(g, "f"), _optr(g, "n")))
    return g


def mex_stokes_dlp_kspace_backward(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P, du, *,
                                   wrt=_WRT_ALL, ctx=None):
    """Gradients of sum(du * mex_stokes_dlp_kspace(...)) (dict, as above)."""
    dev, ps, pt, ff, nn, dd = _adjoint_inputs(psrc, ptar, f, n, du, True)
    ctx = _prepare(ctx, dev)
    g = _grad_outputs(ps.shape[0], pt.shape[0], dev, wrt, True)
    ctx._check(ctx._lib.se2p_dlp_kspace_backward(
        ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0], float(xi), float(eta),
        _ptr(ff), _ptr(nn), int(Mx), int(My), float(Lx), float(Ly), float(w), int(P), _ptr(dd),
        _optr(g, "psrc"), _optr(g, "ptar"), _optr(g, "f"), _optr(g, "n")))
    return g


# -- the reference drivers as single calls --------------------------------------
def _driver(kind, psrc, ptar, f, n, Lx, Ly, P, Nb, tol, M, ctx, same, fixed_grid=False):
    _same_size(psrc, f, "f")
    if kind == 1:
        _same_size(psrc, n, "n")
    dev = _device_of(psrc, ptar, f, n)
    ps = _as_points(psrc, "psrc", dev)
    pt = ps if same else _as_points(ptar, "ptar", dev)
    ff = _as_points(f, "f", dev)
    nn = _as_points(n, "n", dev) if kind == 1 else None
    ctx = _prepare(ctx, dev)
    prm = _capi.Se2pParams()
    prm.P = int(P)
    prm.Nb = float(Nb) if Nb is not None else 0.0
    prm.tol = float(tol)
    prm.M = int(M) if M is not None else 0
    prm.fixed_grid = 1 if fixed_grid else 0
    nt = pt.shape[0]
    u = torch.empty((nt, 2), dtype=torch.float64, device=dev)
    ur = torch.empty_like(u)
    uk = torch.empty_like(u)
    if kind == 0:
        rc = ctx._lib.se2p_slp_ewald(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), nt, _ptr(ff),
                                     float(Lx), float(Ly), ctypes.byref(prm), _ptr(u), _ptr(ur),
                                     _ptr(uk))
    else:
        rc = ctx._lib.se2p_dlp_ewald(ctx.handle, _ptr(ps), ps.shape[0], _ptr(pt), nt, _ptr(ff),
                                     _ptr(nn), float(Lx), float(Ly), ctypes.byref(prm), _ptr(u),
                                     _ptr(ur), _ptr(uk))
    ctx._check(rc)
    if not prm.converged:
        import warnings

        warnings.warn("Couldn't find suitable xi/k: Newton iteration cap reached", RuntimeWarning)
    return u, ur, uk, prm


def autotune(kind, psrc, ptar, f, n, Lx, Ly, *, tol, P=0, ctx=None):
    """Driver parameters for one geometry chosen by timing (se2p_autotune).

    ``kind`` is 'slp' or 'dlp'; points are the caller's (N, 2) arrays
    (unwrapped, as for the drivers); ``ptar is psrc`` means targets are the
    sources.  P = 0 starts from the drivers' support, 24, and lets the
    tuner try narrower supports on oversampled grids.  Returns
    ``(params, info)``: the se2p_params to pass on (``Nb``, ``P`` and ``M``
    set, so ``StokesSLP_ewald_2p(..., P=params.P, Nb=params.Nb,
    M=params.M or None, tol=tol)`` reproduces them) and a dict of the
    candidates timed."""
    k = {"slp": 0, "dlp": 1}[kind]
    same = ptar is psrc
    _same_size(psrc, f, "f")
    if k == 1:
        if n is None:
            raise ValueError("the stresslet needs normals n.")
        _same_size(psrc, n, "n")
    dev = _device_of(psrc, ptar, f, n) if n is not None else _device_of(psrc, ptar, f)
    ps = _as_points(psrc, "psrc", dev)
    pt = ps if same else _as_points(ptar, "ptar", dev)
    ff = _as_points(f, "f", dev)
    nn = _as_points(n, "n", dev) if k == 1 else None
    ctx = _prepare(ctx, dev)
    prm = _capi.Se2pParams()
    prm.P = int(P)
    prm.tol = float(tol)
    info = _capi.Se2pTuneInfo()
    ctx._check(ctx._lib.se2p_autotune(ctx.handle, k, _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0],
                                      _ptr(ff), _ptr(nn) if nn is not None else None, float(Lx),
                                      float(Ly), ctypes.byref(prm), ctypes.byref(info)))
    nc = info.ncand
    out = {"P": info.P, "cached": bool(info.cached), "best": info.best,
           "candidates": [{"Nb": info.Nb[i], "nside_x": info.nside_x[i], "Mx": info.Mx[i],
                           "ms": info.ms[i]} for i in range(nc)],
           "spread_chunk": info.spread_chunk,
           "chunks": [{"chunk": info.chunk[k], "ms": info.chunk_ms[k]} for k in range(3)] if nc else [],
           "supports": [{"P": info.pc_P[i], "Mx": info.pc_Mx[i], "ms": info.pc_ms[i],
                         "dev": info.pc_dev[i]} for i in range(info.npc)]}
    return prm, out


def autotune_multibox(kind, boxes, *, tol, ctx=None):
    """Tunes a multi-box call (se2p_autotune_multibox): each box's cells by
    timing, then the spreading split of the packed call.  ``boxes`` are dicts
    as for the multi-box entries (device (N, 2) tensors "psrc", "ptar", "f",
    "n" for the stresslet, "Lx", "Ly"); their parameter keys (xi, nside_x,
    nside_y, eta, w, Mx, My, P) are overwritten with the tuned values.
    Returns a dict of the timings."""
    k = {"slp": 0, "dlp": 1}[kind]
    if not boxes:
        raise ValueError("no boxes")
    dev = boxes[0]["psrc"].device
    ctx = _prepare(ctx, dev)
    arr = (_capi.Se2pBoxDesc * len(boxes))()
    keep = []
    for i, b in enumerate(boxes):
        ps = _as_points(b["psrc"], "psrc", dev)
        pt = ps if b["ptar"] is b["psrc"] else _as_points(b["ptar"], "ptar", dev)
        ff = _as_points(b["f"], "f", dev)
        nn = _as_points(b["n"], "n", dev) if k == 1 else None
        u = torch.empty((pt.shape[0], 2), dtype=torch.float64, device=dev)
        keep += [ps, pt, ff, nn, u]
        d = arr[i]
        d.psrc, d.ns, d.ptar, d.nt = _ptr(ps), ps.shape[0], _ptr(pt), pt.shape[0]
        d.f, d.n = _ptr(ff), (_ptr(nn) if nn is not None else None)
        d.Lx, d.Ly, d.u = float(b["Lx"]), float(b["Ly"]), _ptr(u)
    info = _capi.Se2pTuneBoxesInfo()
    ctx._check(ctx._lib.se2p_autotune_multibox(ctx.handle, k, arr, len(boxes), float(tol),
                                               ctypes.byref(info)))
    for b, d in zip(boxes, arr):
        b.update(xi=d.xi, nside_x=d.nside_x, nside_y=d.nside_y, eta=d.eta, w=d.w, Mx=d.Mx,
                 My=d.My, P=d.P)
    return {"drivers_ms": info.drivers_ms, "tuned_ms": info.tuned_ms,
            "kept_drivers": bool(info.kept_drivers), "spread_chunk": info.spread_chunk,
            "chunks": [{"chunk": info.chunk[i], "ms": info.chunk_ms[i]} for i in range(3)],
            "smooth_ms": info.smooth_ms, "smooth_grids": info.smooth_grids}


def _cols(*arrays):
    return np.stack([np.asarray(a, dtype=np.float64).ravel() for a in arrays], axis=1)


def StokesSLP_ewald_2p(xsrc, ysrc, xtar, ytar, f1, f2, Lx, Ly, *, P=24, Nb=None, tol=1e-16,
                       M=None, ctx=None, return_params=False, fixed_grid=False):
    """[u1, u2, ur, uk, xi] of PeriodicEwald/src/matlab/StokesSLP_ewald_2p.m:1-138.

    Inputs are 1-D arrays (MATLAB column vectors); ur and uk are (Nt, 2)
    device tensors, u1/u2 device tensors.  ``M`` overrides the grid the
    driver would choose; with ``fixed_grid`` the grid M is fixed and xi and
    the cells follow from it at ``tol`` (se2p_params.fixed_grid; no reference
    counterpart).  ``return_params`` appends the se2p_params used."""
    same = (xsrc is xtar) and (ysrc is ytar)
    ps = _cols(xsrc, ysrc)
    pt = ps if same else _cols(xtar, ytar)
    u, ur, uk, prm = _driver(0, ps, pt, _cols(f1, f2), None, Lx, Ly, P, Nb, tol, M, ctx, same,
                             fixed_grid)
    out = (u[:, 0], u[:, 1], ur, uk, prm.xi)
    return out + (prm,) if return_params else out


def StokesDLP_ewald_2p(xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, Lx, Ly, *, P=24, Nb=None, tol=1e-16,
                       M=None, ctx=None, return_params=False, fixed_grid=False):
    """[u1, u2, ur, uk, xi] of PeriodicEwald/src/matlab/StokesDLP_ewald_2p.m:1-163
    (uk includes the k = 0 mode; u = -(ur + uk)).  Normals are used as given
    (no renormalisation).  ``M`` and ``fixed_grid`` as for StokesSLP_ewald_2p."""
    same = (xsrc is xtar) and (ysrc is ytar)
    ps = _cols(xsrc, ysrc)
    pt = ps if same else _cols(xtar, ytar)
    u, ur, uk, prm = _driver(1, ps, pt, _cols(f1, f2), _cols(n1, n2), Lx, Ly, P, Nb, tol, M, ctx,
                             same, fixed_grid)
    out = (u[:, 0], u[:, 1], ur, uk, prm.xi)
    return out + (prm,) if return_params else out
