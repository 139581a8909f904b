"""ctypes binding to ``libse2p.so`` (the C ABI declared in ``include/se2p.h``).

The library is built in-tree into ``fasttools2d_amd/lib/``.  Loading fails
loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libse2p.so")
CSRC = os.path.join(_HERE, "csrc")

_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64
_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)

# status codes (include/se2p.h)
SE2P_OK = 0
SE2P_ERR_ARG = 1
SE2P_ERR_GRID = 2
SE2P_ERR_P = 3
SE2P_ERR_W = 4
SE2P_ERR_NSIDE = 5
SE2P_ERR_PARAM = 6
SE2P_ERR_NOCONV = 7
SE2P_ERR_HIP = 10
SE2P_ERR_FFT = 11
SE2P_ERR_FROZEN = 12

# every symbol include/se2p.h declares, with (restype, argtypes)
SIGNATURES = {
    "se2p_create": (_I, [_PP, _I]),
    "se2p_destroy": (_I, [_P]),
    "se2p_last_error": (ctypes.c_char_p, [_P]),
    "se2p_set_stream": (_I, [_P, _P]),
    "se2p_get_stream": (_P, [_P]),
    "se2p_set_async": (_I, [_P, _I]),
    "se2p_set_host_pointers": (_I, [_P, _I]),
    "se2p_set_sort_mode": (_I, [_P, _I]),
    "se2p_set_adjoint_passes": (_I, [_P, _I]),
    "se2p_set_keep": (_I, [_P, _I]),
    "se2p_set_spread_chunk": (_I, [_P, _I]),
    "se2p_set_timing": (_I, [_P, _I]),
    "se2p_set_frozen": (_I, [_P, _I]),
    "se2p_stage_times": (_I, [_P, ctypes.POINTER(ctypes.c_float), _I]),
    "se2p_group_begin": (_I, [_P]),
    "se2p_group_end": (_I, [_P]),
    "se2p_slp_real": (_I, [_P, _P, _I64, _P, _I64, _P, _D, _I, _I, _D, _D, _P]),
    "se2p_slp_kspace": (_I, [_P, _P, _I64, _P, _I64, _D, _D, _P, _I, _I, _D, _D, _D, _I, _P]),
    "se2p_dlp_real": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _D, _I, _I, _D, _D, _P]),
    "se2p_dlp_real_noalloc": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _D, _I, _I, _D, _D, _P]),
    "se2p_dlp_kspace": (_I, [_P, _P, _I64, _P, _I64, _D, _D, _P, _P, _I, _I, _D, _D, _D, _I, _P]),
    "se2p_slp_real_batched": (_I, [_P, _P, _I64, _P, _I64, _P, _I, _I64, _D, _I, _I, _D, _D,
                                   _P, _I64]),
    "se2p_slp_kspace_batched": (_I, [_P, _P, _I64, _P, _I64, _D, _D, _P, _I, _I64, _I, _I, _D,
                                     _D, _D, _I, _P, _I64]),
    "se2p_dlp_real_batched": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _I, _I64, _D, _I, _I, _D, _D,
                                   _P, _I64]),
    "se2p_dlp_kspace_batched": (_I, [_P, _P, _I64, _P, _I64, _D, _D, _P, _P, _I, _I64, _I, _I,
                                     _D, _D, _D, _I, _P, _I64]),
    "se2p_slp_real_backward": (_I, [_P, _P, _I64, _P, _I64, _P, _D, _I, _I, _D, _D,
                                    _P, _P, _P, _P]),
    "se2p_slp_kspace_backward": (_I, [_P, _P, _I64, _P, _I64, _D, _D, _P, _I, _I, _D, _D, _D,
                                      _I, _P, _P, _P, _P]),
    "se2p_dlp_real_backward": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _D, _I, _I, _D, _D,
                                    _P, _P, _P, _P, _P]),
    "se2p_dlp_kspace_backward": (_I, [_P, _P, _I64, _P, _I64, _D, _D, _P, _P, _I, _I, _D, _D,
                                      _D, _I, _P, _P, _P, _P, _P]),
    "se2p_count_pairs": (_I, [_P, _P, _I64, _P, _I64, _I, _I, _D, _D,
                              ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "se2p_selftest_e1": (_D, [_D]),
    "se2p_selftest_exp_neg": (_D, [_D]),
    "se2p_selftest_anchor": (None, [_D, _D, _D, _D, _D, _I, _P, _P, _P, _P]),
    "se2p_selftest_cell": (_I, [_D, _D, _D, _D, _I, _I]),
    "se2p_selftest_keysort": (_I, [_P, _P, _I64, _I, _P, _P]),
}

SE2P_HOST_ONLY = -1
SE2P_NOALLOC_SCRATCH_BYTES = 32 * 1024 * 1024

class Se2pParams(ctypes.Structure):
    """se2p_params of include/se2p.h."""
    _fields_ = [("P", ctypes.c_int), ("Nb", ctypes.c_double), ("tol", ctypes.c_double),
                ("M", ctypes.c_int), ("nside_x", ctypes.c_int), ("nside_y", ctypes.c_int),
                ("Mx", ctypes.c_int), ("My", ctypes.c_int), ("kinf", ctypes.c_int),
                ("converged", ctypes.c_int), ("rc", ctypes.c_double), ("xi", ctypes.c_double),
                ("eta", ctypes.c_double), ("w", ctypes.c_double), ("fixed_grid", ctypes.c_int)]


SIGNATURES.update({
    "se2p_select_params": (_I, [_I, _I64, _I64, _D, _D, _D, ctypes.POINTER(Se2pParams)]),
    "se2p_slp_ewald": (_I, [_P, _P, _I64, _P, _I64, _P, _D, _D, ctypes.POINTER(Se2pParams),
                            _P, _P, _P]),
    "se2p_dlp_ewald": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _D, _D, ctypes.POINTER(Se2pParams),
                            _P, _P, _P]),
})

class Se2pTuneInfo(ctypes.Structure):
    """se2p_tune_info of include/se2p.h."""
    _fields_ = [("ncand", ctypes.c_int), ("best", ctypes.c_int), ("cached", ctypes.c_int),
                ("P", ctypes.c_int), ("Nb", ctypes.c_double * 16), ("nside_x", ctypes.c_int * 16),
                ("Mx", ctypes.c_int * 16), ("ms", ctypes.c_double * 16),
                ("spread_chunk", ctypes.c_int), ("chunk", ctypes.c_int * 3),
                ("chunk_ms", ctypes.c_double * 3), ("npc", ctypes.c_int),
                ("pc_P", ctypes.c_int * 6), ("pc_Mx", ctypes.c_int * 6),
                ("pc_ms", ctypes.c_double * 6), ("pc_dev", ctypes.c_double * 6)]


class Se2pTuneBoxesInfo(ctypes.Structure):
    """se2p_tune_boxes_info of include/se2p.h."""
    _fields_ = [("drivers_ms", ctypes.c_double), ("tuned_ms", ctypes.c_double),
                ("kept_drivers", ctypes.c_int), ("spread_chunk", ctypes.c_int),
                ("chunk", ctypes.c_int * 3), ("chunk_ms", ctypes.c_double * 3),
                ("smooth_ms", ctypes.c_double), ("smooth_grids", ctypes.c_int)]


SIGNATURES.update({
    "se2p_autotune": (_I, [_P, _I, _P, _I64, _P, _I64, _P, _P, _D, _D, ctypes.POINTER(Se2pParams),
                           ctypes.POINTER(Se2pTuneInfo)]),
})

class Se2pBoxDesc(ctypes.Structure):
    """se2p_box_desc of include/se2p.h."""
    _fields_ = [("psrc", ctypes.c_void_p), ("ns", ctypes.c_int64),
                ("ptar", ctypes.c_void_p), ("nt", ctypes.c_int64),
                ("f", ctypes.c_void_p), ("n", ctypes.c_void_p),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("xi", ctypes.c_double),
                ("nside_x", ctypes.c_int), ("nside_y", ctypes.c_int),
                ("eta", ctypes.c_double), ("w", ctypes.c_double),
                ("Mx", ctypes.c_int), ("My", ctypes.c_int), ("P", ctypes.c_int),
                ("u", ctypes.c_void_p)]


SIGNATURES.update({
    name: (_I, [_P, ctypes.POINTER(Se2pBoxDesc), _I])
    for name in ("se2p_slp_real_multibox", "se2p_slp_kspace_multibox",
                 "se2p_dlp_real_multibox", "se2p_dlp_kspace_multibox")
})
SIGNATURES.update({
    "se2p_autotune_multibox": (_I, [_P, _I, ctypes.POINTER(Se2pBoxDesc), _I, _D,
                                    ctypes.POINTER(Se2pTuneBoxesInfo)]),
})

STAGES = ("sort", "real", "spread", "fft_fwd", "filter", "fft_inv", "interp", "pair")

_lib = None


class Se2pError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"se2p error {code}: {msg}")
        self.code = code
        self.msg = msg


def load():
    """Load libse2p.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C {CSRC}` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def build() -> None:
    import subprocess

    subprocess.run(["make", "-s", "-j8", "-C", CSRC], check=True)
