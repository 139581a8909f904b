"""Argument checks of se2p_autotune on a host-only context (CPU)."""
import ctypes

import numpy as np

from fasttools2d_amd import _capi


def test_autotune_argument_checks():
    lib = _capi.load()
    h = ctypes.c_void_p()
    assert lib.se2p_create(ctypes.byref(h), _capi.SE2P_HOST_ONLY) == 0
    try:
        x = np.zeros((4, 2))
        ptr = x.ctypes.data_as(ctypes.c_void_p)

        def call(kind=0, tol=1e-10, P=0, M=0, Lx=1.0, ns=4):
            prm = _capi.Se2pParams()
            prm.tol, prm.P, prm.M = tol, P, M
            return lib.se2p_autotune(h, kind, ptr, ns, ptr, 4, ptr, ptr, Lx, 1.0,
                                     ctypes.byref(prm), None)

        assert call(tol=0.0) == _capi.SE2P_ERR_PARAM
        assert call(tol=2.0) == _capi.SE2P_ERR_PARAM
        assert call(P=5) == _capi.SE2P_ERR_P
        assert call(M=64) == _capi.SE2P_ERR_ARG
        assert call(kind=2) == _capi.SE2P_ERR_ARG
        assert call(Lx=-1.0) == _capi.SE2P_ERR_PARAM
        assert call(ns=-1) == _capi.SE2P_ERR_ARG
        # valid arguments reach the device step, which a host-only context refuses
        assert call() == _capi.SE2P_ERR_HIP
        assert b"host-only" in lib.se2p_last_error(h)
    finally:
        lib.se2p_destroy(h)


def test_spread_chunk_argument_checks():
    lib = _capi.load()
    h = ctypes.c_void_p()
    assert lib.se2p_create(ctypes.byref(h), _capi.SE2P_HOST_ONLY) == 0
    try:
        for ok in (0, 64, 128, 256, 512, 2048):
            assert lib.se2p_set_spread_chunk(h, ok) == 0, ok
        for bad in (-16, 16, 48, 100, 2064, 4096):
            assert lib.se2p_set_spread_chunk(h, bad) == _capi.SE2P_ERR_ARG, bad
            assert b"spread chunk" in lib.se2p_last_error(h)
        assert lib.se2p_set_spread_chunk(None, 0) == _capi.SE2P_ERR_ARG
    finally:
        lib.se2p_destroy(h)
