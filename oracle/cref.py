"""ctypes binding to the C oracle ``oracle/_build/libse2p_oracle.so``.

Test infrastructure only (see ``oracle/__init__.py``).  Array conventions
follow the reference MEX boundary: points, densities and normals are ``(N, 2)``
float64 C-contiguous arrays, i.e. MATLAB's 2xN column-major layout
``[x0, y0, x1, y1, ...]``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# builds of the same source by instruction set (oracle/Makefile); the tests use
# the portable x86-64-v2 build, the CPU-baseline leg the best the host runs
_BUILDS = {"x86-64-v2": "libse2p_oracle.so", "x86-64-v3": "libse2p_oracle_v3.so",
           "x86-64-v4": "libse2p_oracle_v4.so"}
_LIB_PATH = os.path.join(_HERE, "_build", _BUILDS["x86-64-v2"])
_lib = None


def host_isa() -> str:
    """The highest x86-64 level of this host's CPU (from /proc/cpuinfo):
    x86-64-v4 (AVX-512 F/BW/CD/DQ/VL), x86-64-v3 (AVX2, FMA, BMI1/2, MOVBE,
    F16C) or x86-64-v2."""
    try:
        with open("/proc/cpuinfo") as fh:
            flags = set()
            for line in fh:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        return "x86-64-v2"
    v3 = {"avx2", "fma", "bmi1", "bmi2", "movbe", "f16c", "abm"}
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
    if v3 <= flags and v4 <= flags:
        return "x86-64-v4"
    if v3 <= flags:
        return "x86-64-v3"
    return "x86-64-v2"


def use_isa(isa: str) -> None:
    """Load the build for instruction set `isa` (before the first call)."""
    global _LIB_PATH
    path = os.path.join(_HERE, "_build", _BUILDS[isa])
    if _lib is not None and path != _LIB_PATH:
        raise RuntimeError("the oracle is already loaded; choose the build before the first call")
    _LIB_PATH = path

_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64
_P = ctypes.c_void_p


def build() -> None:
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        if not os.path.exists(_LIB_PATH):
            raise ImportError(f"{_LIB_PATH} is missing: build it with `make -C {_HERE}`")
        L = ctypes.CDLL(_LIB_PATH)
        L.or_e1.restype = _D
        L.or_e1.argtypes = [_D]
        L.or_find_closest_node.restype = None
        L.or_find_closest_node.argtypes = [_D, _D, _D, _D, _D, _I, _P, _P, _P, _P]
        L.or_cell_lists.restype = None
        L.or_cell_lists.argtypes = [_P, _I64, _D, _D, _I, _I, _P, _P]
        L.or_slp_real.restype = _I
        L.or_slp_real.argtypes = [_P, _I64, _P, _I64, _P, _D, _I, _I, _D, _D, _P]
        L.or_dlp_real.restype = _I
        L.or_dlp_real.argtypes = [_P, _I64, _P, _I64, _P, _P, _D, _I, _I, _D, _D, _P]
        L.or_slp_kspace.restype = _I
        L.or_slp_kspace.argtypes = [_P, _I64, _P, _I64, _D, _D, _P, _I, _I, _D, _D, _D, _I, _P]
        L.or_dlp_kspace.restype = _I
        L.or_dlp_kspace.argtypes = [_P, _I64, _P, _I64, _D, _D, _P, _P, _I, _I, _D, _D, _D, _I, _P]
        L.or_slp_real_adj.restype = _I
        L.or_slp_real_adj.argtypes = [_P, _I64, _P, _I64, _P, _D, _I, _I, _D, _D, _P, _P, _P, _P]
        L.or_dlp_real_adj.restype = _I
        L.or_dlp_real_adj.argtypes = [_P, _I64, _P, _I64, _P, _P, _D, _I, _I, _D, _D,
                                      _P, _P, _P, _P, _P]
        L.or_slp_kspace_adj.restype = _I
        L.or_slp_kspace_adj.argtypes = [_P, _I64, _P, _I64, _D, _D, _P, _I, _I, _D, _D, _D, _I,
                                        _P, _P, _P, _P]
        L.or_dlp_kspace_adj.restype = _I
        L.or_dlp_kspace_adj.argtypes = [_P, _I64, _P, _I64, _D, _D, _P, _P, _I, _I, _D, _D, _D,
                                        _I, _P, _P, _P, _P, _P]
        L.or_fft2.restype = None
        L.or_fft2.argtypes = [_P, _P, _I, _I, _I]
        L.or_set_threads.restype = None
        L.or_set_threads.argtypes = [_I]
        L.or_get_threads.restype = _I
        L.or_get_threads.argtypes = []
        _lib = L
    return _lib


def _pts(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError("expected an (N, 2) array (MATLAB 2xN layout)")
    return a


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def e1(x: float) -> float:
    return lib().or_e1(float(x))


def find_closest_node(x, y, Lx, Ly, h, P):
    mx, my = _I(), _I()
    px, py = _D(), _D()
    lib().or_find_closest_node(float(x), float(y), float(Lx), float(Ly), float(h), int(P),
                               ctypes.byref(mx), ctypes.byref(my),
                               ctypes.byref(px), ctypes.byref(py))
    return mx.value, my.value, px.value, py.value


def cell_lists(p, Lx, Ly, nsx, nsy):
    p = _pts(p)
    perm = np.zeros(max(len(p), 1), dtype=np.int64)
    off = np.zeros(nsx * nsy + 1, dtype=np.int64)
    lib().or_cell_lists(_ptr(p), len(p), float(Lx), float(Ly), int(nsx), int(nsy),
                        _ptr(perm), _ptr(off))
    return perm[: len(p)], off


def slp_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly):
    """mex_stokes_slp_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly)."""
    psrc, ptar, f = _pts(psrc), _pts(ptar), _pts(f)
    u = np.zeros((len(ptar), 2))
    lib().or_slp_real(_ptr(psrc), len(psrc), _ptr(ptar), len(ptar), _ptr(f), float(xi),
                      int(nside_x), int(nside_y), float(Lx), float(Ly), _ptr(u))
    return u


def dlp_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly):
    """mex_stokes_dlp_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly)."""
    psrc, ptar, f, n = _pts(psrc), _pts(ptar), _pts(f), _pts(n)
    u = np.zeros((len(ptar), 2))
    lib().or_dlp_real(_ptr(psrc), len(psrc), _ptr(ptar), len(ptar), _ptr(f), _ptr(n),
                      float(xi), int(nside_x), int(nside_y), float(Lx), float(Ly), _ptr(u))
    return u


def slp_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P):
    """mex_stokes_slp_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P)."""
    psrc, ptar, f = _pts(psrc), _pts(ptar), _pts(f)
    u = np.zeros((len(ptar), 2))
    lib().or_slp_kspace(_ptr(psrc), len(psrc), _ptr(ptar), len(ptar), float(xi), float(eta),
                        _ptr(f), int(Mx), int(My), float(Lx), float(Ly), float(w), int(P),
                        _ptr(u))
    return u


def dlp_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P):
    """mex_stokes_dlp_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P)."""
    psrc, ptar, f, n = _pts(psrc), _pts(ptar), _pts(f), _pts(n)
    u = np.zeros((len(ptar), 2))
    lib().or_dlp_kspace(_ptr(psrc), len(psrc), _ptr(ptar), len(ptar), float(xi), float(eta),
                        _ptr(f), _ptr(n), int(Mx), int(My), float(Lx), float(Ly), float(w),
                        int(P), _ptr(u))
    return u


def _grads(ns, nt, dlp):
    out = {"dpsrc": np.zeros((ns, 2)), "dptar": np.zeros((nt, 2)), "df": np.zeros((ns, 2))}
    if dlp:
        out["dn"] = np.zeros((ns, 2))
    return out


def slp_real_adj(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, du):
    """Gradients of L = sum du . slp_real(...) w.r.t. psrc, ptar and f (dict)."""
    psrc, ptar, f, du = _pts(psrc), _pts(ptar), _pts(f), _pts(du)
    o = _grads(len(psrc), len(ptar), False)
    lib().or_slp_real_adj(_ptr(psrc), len(psrc), _ptr(ptar), len(ptar), _ptr(f), float(xi),
                          int(nside_x), int(nside_y), float(Lx), float(Ly), _ptr(du),
                          _ptr(o["dpsrc"]), _ptr(o["dptar"]), _ptr(o["df"]))
    return o


def dlp_real_adj(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, du):
    """Gradients of L = sum du . dlp_real(...) w.r.t. psrc, ptar, f and n (dict)."""
    psrc, ptar, f, n, du = _pts(psrc), _pts(ptar), _pts(f), _pts(n), _pts(du)
    o = _grads(len(psrc), len(ptar), True)
    lib().or_dlp_real_adj(_ptr(psrc), len(psrc), _ptr(ptar), len(ptar), _ptr(f), _ptr(n),
                          float(xi), int(nside_x), int(nside_y), float(Lx), float(Ly), _ptr(du),
                          _ptr(o["dpsrc"]), _ptr(o["dptar"]), _ptr(o["df"]), _ptr(o["dn"]))
    return o


def slp_kspace_adj(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P, du):
    """Gradients of L = sum du . slp_kspace(...) w.r.t. psrc, ptar and f (dict)."""
    psrc, ptar, f, du = _pts(psrc), _pts(ptar), _pts(f), _pts(du)
    o = _grads(len(psrc), len(ptar), False)
    lib().or_slp_kspace_adj(_ptr(psrc), len(psrc), _ptr(ptar), len(ptar), float(xi), float(eta),
                            _ptr(f), int(Mx), int(My), float(Lx), float(Ly), float(w), int(P),
                            _ptr(du), _ptr(o["dpsrc"]), _ptr(o["dptar"]), _ptr(o["df"]))
    return o


def dlp_kspace_adj(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P, du):
    """Gradients of L = sum du . dlp_kspace(...) w.r.t. psrc, ptar, f and n (dict)."""
    psrc, ptar, f, n, du = _pts(psrc), _pts(ptar), _pts(f), _pts(n), _pts(du)
    o = _grads(len(psrc), len(ptar), True)
    lib().or_dlp_kspace_adj(_ptr(psrc), len(psrc), _ptr(ptar), len(ptar), float(xi), float(eta),
                            _ptr(f), _ptr(n), int(Mx), int(My), float(Lx), float(Ly), float(w),
                            int(P), _ptr(du), _ptr(o["dpsrc"]), _ptr(o["dptar"]), _ptr(o["df"]),
                            _ptr(o["dn"]))
    return o


def fft2(z: np.ndarray, inverse: bool = False) -> np.ndarray:
    """2D DFT of an (Mx, My) complex array whose memory is the reference's
    My x Mx column-major grid (index x*My + y)."""
    z = np.asarray(z, dtype=np.complex128)
    Mx, My = z.shape
    re = np.ascontiguousarray(z.real)
    im = np.ascontiguousarray(z.imag)
    lib().or_fft2(_ptr(re), _ptr(im), Mx, My, 1 if inverse else -1)
    return re + 1j * im


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return lib().or_get_threads()
