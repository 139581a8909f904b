"""Host-side logic without a GPU: parameter selection of the reference drivers,
the C ABI's exports and its argument validation, and the product's host
builds of the device E1 and cell-index arithmetic."""
import ctypes
import math
import os
import re

import mpmath
import numpy as np
import pytest

from fasttools2d_amd import _capi
from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_matlab_rat_and_round():
    assert Pm.matlab_rat(1.0) == (1, 1)
    assert Pm.matlab_rat(0.5) == (1, 2)
    assert Pm.matlab_rat(2.0) == (2, 1)
    assert Pm.matlab_rat(1.5) == (3, 2)
    assert Pm.matlab_rat(4 / 3) == (4, 3)
    n, d = Pm.matlab_rat(math.pi)
    assert abs(n / d - math.pi) < 1e-6 * math.pi
    assert Pm.matlab_round(2.5) == 3 and Pm.matlab_round(-2.5) == -3 and Pm.matlab_round(0.49) == 0


def test_wrap():
    x = np.array([0.0, 0.49, 0.5, 0.99, 1.0, -0.5, -0.51])
    assert np.allclose(Pm.wrap(x, 1.0), [0.0, 0.49, -0.5, -0.01, 0.0, -0.5, 0.49])


def test_driver_parameters_spot_values():
    # 1000 uniform points, targets = sources, tol 1e-10 (unit box)
    w = W.uniform_box(1000, seed=1)
    p = Pm.select_params("slp", 1000, 1000, w.f, 1.0, 1.0, tol=1e-10)
    assert (p.nside_x, p.nside_y) == (3, 3)
    assert abs(p.xi - 15.04) < 0.01 and p.kinf == 24 and p.Mx == 48
    q = Pm.select_params("dlp", 1000, 1000, w.f, 1.0, 1.0, tol=1e-10)
    assert abs(q.xi - 16.15) < 0.02 and q.kinf == 28 and q.Mx == 56
    # the error estimates are at the tolerance after the Newton solve
    Q = float(np.sum(w.f ** 2)) + 1
    f_slp = math.sqrt(Q * math.pi / (4 * p.xi)) * math.exp(-p.xi ** 2 * p.rc ** 2)
    assert abs(f_slp - 1e-10) <= 1e-10
    f_dlp = math.exp(-q.xi ** 2 * q.rc ** 2) * q.xi * q.rc * math.sqrt(2 * math.pi * Q)
    assert abs(f_dlp - 1e-10) <= 1e-10


def test_grid_override_keeps_w_and_eta_consistent():
    w = W.uniform_box(200, seed=2)
    p = Pm.select_params("slp", 200, 200, w.f, 1.0, 1.0, tol=1e-10, M=64)
    assert p.Mx == 64 and p.w == 24 * 1.0 / 64 / 2
    m = 0.95 * math.sqrt(math.pi * 24)
    assert abs(p.eta - (2 * p.xi * p.w / m) ** 2) < 1e-15


def test_ellipse_workload_geometry():
    w = W.ellipses(n_ell=20, n_nodes=100, seed=3)
    assert w.psrc.shape == (2000, 2) and np.all(np.abs(w.psrc) <= 0.5)
    assert np.allclose(np.linalg.norm(w.n, axis=1), 1.0)
    c = w.meta["centres"]
    # outward normals: (x - centre).n > 0
    xu = w.psrc_unwrapped.reshape(20, 100, 2)
    nn = w.n.reshape(20, 100, 2)
    assert np.all(((xu - c[:, None, :]) * nn).sum(axis=2) > 0)
    # bounding circles at least the gap apart (periodic)
    R = w.meta["semi_axes"].max(axis=1)
    for i in range(20):
        d = c - c[i]
        d -= np.round(d)
        dist = np.sqrt((d ** 2).sum(axis=1))
        dist[i] = np.inf
        assert np.all(dist >= R + R[i] + 0.005 - 1e-12)


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "se2p.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(se2p_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    decl = _declared_symbols()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(lib, name), name
        assert name in _capi.SIGNATURES, name


def test_status_codes_match_header():
    txt = open(os.path.join(ROOT, "include", "se2p.h")).read()
    for name in ("SE2P_OK", "SE2P_ERR_ARG", "SE2P_ERR_GRID", "SE2P_ERR_P", "SE2P_ERR_W",
                 "SE2P_ERR_NSIDE", "SE2P_ERR_PARAM", "SE2P_ERR_NOCONV", "SE2P_ERR_HIP",
                 "SE2P_ERR_FFT"):
        m = re.search(r"#define\s+" + name + r"\s+(\d+)", txt)
        assert m and int(m.group(1)) == getattr(_capi, name), name


@pytest.fixture()
def hctx():
    lib = _capi.load()
    h = ctypes.c_void_p()
    assert lib.se2p_create(ctypes.byref(h), _capi.SE2P_HOST_ONLY) == 0
    yield lib, h
    lib.se2p_destroy(h)


def _fake(n):
    # host pointers are never dereferenced by the validation paths
    return ctypes.c_void_p(0x1000) if n else None


def test_kspace_validation_codes(hctx):
    lib, h = hctx
    P, M, L = 24, 64, 1.0
    w = P * L / M / 2
    x, u = _fake(1), _fake(1)

    def call(**kw):
        a = dict(xi=10.0, eta=0.5, Mx=M, My=M, Lx=L, Ly=L, w=w, P=P)
        a.update(kw)
        return lib.se2p_slp_kspace(h, x, 4, x, 4, a["xi"], a["eta"], x, a["Mx"], a["My"],
                                   a["Lx"], a["Ly"], a["w"], a["P"], u)

    assert call(Mx=63) == _capi.SE2P_ERR_GRID          # odd grid
    assert call(Mx=64, My=32) == _capi.SE2P_ERR_GRID   # Lx/Mx != Ly/My
    assert call(Mx=12, My=12, w=P * L / 12 / 2) == _capi.SE2P_ERR_GRID   # M < P/2 + 2
    assert call(P=23) == _capi.SE2P_ERR_P
    assert call(w=w * 1.01) == _capi.SE2P_ERR_W
    assert call(xi=-1.0) == _capi.SE2P_ERR_PARAM
    assert call(eta=0.0) == _capi.SE2P_ERR_PARAM
    # valid arguments reach the device step, which a host-only context refuses
    assert call() == _capi.SE2P_ERR_HIP
    assert lib.se2p_dlp_kspace(h, x, -1, x, 4, 10.0, 0.5, x, x, M, M, L, L, w, P, u) == _capi.SE2P_ERR_ARG
    assert lib.se2p_dlp_kspace(h, x, 4, x, 4, 10.0, 0.5, x, None, M, M, L, L, w, P, u) == _capi.SE2P_ERR_ARG


def test_sort_mode_option(hctx):
    lib, h = hctx
    for mode in (0, 1, 2, 3):
        assert lib.se2p_set_sort_mode(h, mode) == 0
    assert lib.se2p_set_sort_mode(h, 4) == _capi.SE2P_ERR_ARG
    assert b"sort mode" in lib.se2p_last_error(h)
    assert lib.se2p_set_sort_mode(None, 0) == _capi.SE2P_ERR_ARG


def test_real_validation_codes(hctx):
    lib, h = hctx
    x = _fake(1)
    assert lib.se2p_slp_real(h, x, 4, x, 4, x, 10.0, 0, 3, 1.0, 1.0, x) == _capi.SE2P_ERR_NSIDE
    assert lib.se2p_dlp_real(h, x, 4, x, 4, x, x, 10.0, 3, 3, 1.0, -1.0, x) == _capi.SE2P_ERR_PARAM
    assert lib.se2p_dlp_real(h, None, 4, x, 4, x, x, 10.0, 3, 3, 1.0, 1.0, x) == _capi.SE2P_ERR_ARG
    assert lib.se2p_slp_real(h, x, 4, x, 4, x, 10.0, 3, 3, 1.0, 1.0, x) == _capi.SE2P_ERR_HIP


def test_product_e1_against_mpmath():
    """The kernels' piecewise E1 (host build of the same code) to a few ulp."""
    lib = _capi.load()
    xs = np.concatenate([np.logspace(-14, -0.01, 80), [0.5, 1.0, 1.25, 2.0, 600.0],
                         np.linspace(1.001, 80, 300)])
    worst = 0.0
    for x in xs:
        ref = float(mpmath.e1(mpmath.mpf(float(x))))
        worst = max(worst, abs(lib.se2p_selftest_e1(float(x)) - ref) / ref)
    assert worst < 2e-15, worst
    assert lib.se2p_selftest_e1(2000.0) == 0.0


def test_product_exp_against_mpmath():
    """The real-space kernels' table-driven exp(-x) (host build of the same
    code) over the range the pair sums use (x = xi^2 r^2 below about 50), the
    table seams (x = k ln2/64 and half-way) and large x."""
    lib = _capi.load()
    ln2_64 = float(mpmath.log(2) / 64)
    xs = np.concatenate([[0.0, 1e-300, 1e-17, 1e-9], np.linspace(0, 60, 3001),
                         ln2_64 * np.arange(0, 4000, 7), ln2_64 * (np.arange(0, 4000, 7) + 0.5),
                         [100.0, 300.0, 700.0]])
    worst = 0.0
    for x in xs:
        ref = mpmath.exp(-mpmath.mpf(float(x)))
        got = lib.se2p_selftest_exp_neg(float(x))
        worst = max(worst, float(abs(got - ref) / ref))
    assert worst < 1.5e-15, worst


def test_product_cell_index_matches_oracle_cells():
    from oracle import cref
    lib = _capi.load()
    rng = np.random.default_rng(4)
    p = Pm.wrap(rng.uniform(0, 1, (300, 2)), 1.0)
    perm, off = cref.cell_lists(p, 1.0, 1.0, 7, 5)
    cells = np.array([lib.se2p_selftest_cell(x, y, 1.0, 1.0, 7, 5) for x, y in p])
    for c in range(35):
        assert set(perm[off[c]:off[c + 1]]) == set(np.nonzero(cells == c)[0])


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_c_parameter_selection_matches_host_restatement(kind):
    """se2p_select_params (C++) against fasttools2d_amd.params (Python): both
    restate the drivers' Newton solves; they must pick the same cells and
    grids and the same xi to rounding."""
    lib = _capi.load()
    for (N, Lx, Ly, tol, seed) in ((1000, 1.0, 1.0, 1e-10, 1), (700, 1.0, 2.0, 1e-12, 2),
                                   (5000, 3.0, 1.0, 1e-8, 3)):
        rng = np.random.default_rng(seed)
        f = rng.standard_normal((N, 2))
        Q = float(np.sum(f * f)) + 1
        ref = Pm.select_params(kind, N, N, f, Lx, Ly, tol=tol)
        prm = _capi.Se2pParams()
        prm.tol = tol
        assert lib.se2p_select_params(0 if kind == "slp" else 1, N, N, Q, Lx, Ly, ctypes.byref(prm)) == 0
        assert (prm.nside_x, prm.nside_y, prm.Mx, prm.My, prm.kinf) == \
            (ref.nside_x, ref.nside_y, ref.Mx, ref.My, ref.kinf)
        assert abs(prm.xi - ref.xi) <= 1e-12 * ref.xi
        assert abs(prm.eta - ref.eta) <= 1e-12 * ref.eta and prm.converged == 1


def test_backward_validation_codes(hctx):
    lib, h = hctx
    x = _fake(1)
    P, M, L = 24, 64, 1.0
    w = P * L / M / 2
    # du is required; outputs may be null
    assert lib.se2p_slp_real_backward(h, x, 4, x, 4, x, 10.0, 3, 3, L, L, None, x, x, x) == _capi.SE2P_ERR_ARG
    assert lib.se2p_dlp_real_backward(h, x, 4, x, 4, x, None, 10.0, 3, 3, L, L, x, x, x, x, x) == _capi.SE2P_ERR_ARG
    assert lib.se2p_slp_real_backward(h, x, 4, x, 4, x, 10.0, 0, 3, L, L, x, None, None, None) == _capi.SE2P_ERR_NSIDE
    assert lib.se2p_slp_kspace_backward(h, x, 4, x, 4, 10.0, 0.5, x, 63, M, L, L, w, P, x, x, x, x) == _capi.SE2P_ERR_GRID
    assert lib.se2p_dlp_kspace_backward(h, x, 4, x, 4, 10.0, 0.5, x, x, M, M, L, L, w * 1.01, P, x, x, x, x, x) == _capi.SE2P_ERR_W
    assert lib.se2p_dlp_kspace_backward(h, x, -2, x, 4, 10.0, 0.5, x, x, M, M, L, L, w, P, x, x, x, x, x) == _capi.SE2P_ERR_ARG
    # valid arguments, null outputs allowed: reach the device step
    assert lib.se2p_slp_real_backward(h, x, 4, x, 4, x, 10.0, 3, 3, L, L, x, None, x, None) == _capi.SE2P_ERR_HIP
    assert lib.se2p_dlp_kspace_backward(h, x, 4, x, 4, 10.0, 0.5, x, x, M, M, L, L, w, P, x, x, None, None, x) == _capi.SE2P_ERR_HIP


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of se2p_params, se2p_box_desc and the tuner info
    structs have the C layout."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "se2p.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(se2p_params), sizeof(se2p_box_desc),'
        ' offsetof(se2p_box_desc, Lx), offsetof(se2p_box_desc, Mx), offsetof(se2p_box_desc, u),'
        ' sizeof(se2p_tune_info), offsetof(se2p_tune_info, pc_dev), sizeof(se2p_tune_boxes_info),'
        ' offsetof(se2p_tune_boxes_info, smooth_grids));return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(_capi.Se2pParams), ctypes.sizeof(_capi.Se2pBoxDesc),
            _capi.Se2pBoxDesc.Lx.offset, _capi.Se2pBoxDesc.Mx.offset, _capi.Se2pBoxDesc.u.offset,
            ctypes.sizeof(_capi.Se2pTuneInfo), _capi.Se2pTuneInfo.pc_dev.offset,
            ctypes.sizeof(_capi.Se2pTuneBoxesInfo), _capi.Se2pTuneBoxesInfo.smooth_grids.offset]
    assert got == want, (got, want)


def test_cell_count_overflow_rejected_by_every_real_entry(hctx):
    """nside_x * nside_y * 16 (cell and 4x4 sub-cell keys) must fit in a
    32-bit int; every real-space entry refuses more cells before any device
    work, the backward entries included."""
    lib, h = hctx
    x = _fake(1)
    L = 1.0
    for ns in (16384, 11586):
        assert lib.se2p_slp_real(h, x, 4, x, 4, x, 10.0, ns, ns, L, L, x) == _capi.SE2P_ERR_NSIDE
        assert lib.se2p_dlp_real(h, x, 4, x, 4, x, x, 10.0, ns, ns, L, L, x) == _capi.SE2P_ERR_NSIDE
        assert lib.se2p_slp_real_backward(h, x, 4, x, 3, x, 10.0, ns, ns, L, L, x, x, x, x) == \
            _capi.SE2P_ERR_NSIDE
        assert lib.se2p_dlp_real_backward(h, x, 4, x, 3, x, x, 10.0, ns, ns, L, L, x, x, x, x, x) == \
            _capi.SE2P_ERR_NSIDE
        assert lib.se2p_dlp_real_noalloc(h, x, 4, x, 4, x, x, 10.0, ns, ns, L, L, x) == \
            _capi.SE2P_ERR_NSIDE
        a, b = ctypes.c_int64(), ctypes.c_int64()
        assert lib.se2p_count_pairs(h, x, 4, x, 4, ns, ns, L, L, ctypes.byref(a), ctypes.byref(b)) == \
            _capi.SE2P_ERR_NSIDE
        assert b"32-bit" in lib.se2p_last_error(h)
    # the largest square cell grid that fits reaches the device step
    assert lib.se2p_slp_real_backward(h, x, 4, x, 3, x, 10.0, 11585, 11585, L, L, x, x, x, x) == \
        _capi.SE2P_ERR_HIP


def test_noalloc_entry_validation(hctx):
    """se2p_dlp_real_noalloc: the fixed scratch's size limit and the refusal of
    host-pointer mode are checked before any device work."""
    lib, h = hctx
    x = _fake(1)
    assert lib.se2p_dlp_real_noalloc(h, x, 400000, x, 400000, x, x, 100.0, 22, 22, 1.0, 1.0, x) == \
        _capi.SE2P_ERR_ARG
    msg = lib.se2p_last_error(h).decode()
    need = int(re.search(r"need (\d+) bytes", msg).group(1))
    assert need > _capi.SE2P_NOALLOC_SCRATCH_BYTES and "fixed scratch holds 33554432" in msg
    # the benchmark's size fits (and then needs a device)
    assert lib.se2p_dlp_real_noalloc(h, x, 100000, x, 100000, x, x, 100.0, 22, 22, 1.0, 1.0, x) == \
        _capi.SE2P_ERR_HIP
    assert lib.se2p_dlp_real_noalloc(h, x, 0, x, 0, None, None, 100.0, 22, 22, 1.0, 1.0, x) == \
        _capi.SE2P_ERR_HIP
    assert lib.se2p_set_host_pointers(h, 1) == 0
    assert lib.se2p_dlp_real_noalloc(h, x, 10, x, 10, x, x, 10.0, 1, 1, 1.0, 1.0, x) == \
        _capi.SE2P_ERR_ARG
    assert b"device pointers only" in lib.se2p_last_error(h)


def test_frozen_flag_setter(hctx):
    lib, h = hctx
    assert lib.se2p_set_frozen(h, 1) == 0 and lib.se2p_set_frozen(h, 0) == 0
    assert lib.se2p_set_frozen(None, 1) == _capi.SE2P_ERR_ARG


def test_group_entry_validation(hctx):
    """se2p_group_begin / se2p_group_end on a context without a device, and
    their null-context and not-open errors (the GPU behaviour is in
    tests/test_gpu_group.py)."""
    lib, h = hctx
    assert lib.se2p_group_end(h) == _capi.SE2P_ERR_ARG
    assert b"no group is open" in lib.se2p_last_error(h)
    assert lib.se2p_group_begin(h) == _capi.SE2P_ERR_HIP
    assert lib.se2p_group_end(h) == _capi.SE2P_ERR_ARG   # the failed begin opened nothing
    assert lib.se2p_group_begin(None) == _capi.SE2P_ERR_ARG
    assert lib.se2p_group_end(None) == _capi.SE2P_ERR_ARG


def test_python_out_argument_checked():
    """out= is written through its data pointer by the library, so anything but
    a contiguous float64 (Nt, 2) tensor on the inputs' device is refused."""
    import torch

    from fasttools2d_amd import periodic_ewald as pe

    dev = torch.device("cpu")
    good = torch.empty((5, 2), dtype=torch.float64)
    assert pe._output(good, 5, dev) is good
    assert pe._output(None, 5, dev).shape == (5, 2)
    bad = [torch.empty((5, 2), dtype=torch.float32), torch.empty((4, 2), dtype=torch.float64),
           torch.empty((2, 5), dtype=torch.float64).t(), torch.empty((10,), dtype=torch.float64),
           np.zeros((5, 2))]
    for b in bad:
        with pytest.raises(ValueError):
            pe._output(b, 5, dev)
    with pytest.raises(ValueError, match="out must be on"):
        pe._output(good, 5, torch.device("meta"))


def test_python_driver_and_autotune_check_density_sizes():
    """The drivers and the autotuner refuse densities or normals whose length
    differs from the sources' before anything reaches the device."""
    from fasttools2d_amd import periodic_ewald as pe

    x = np.zeros(10)
    with pytest.raises(ValueError, match="psrc and f must be the same size"):
        pe.StokesSLP_ewald_2p(x, x, x, x, np.zeros(9), np.zeros(9), 1.0, 1.0)
    with pytest.raises(ValueError, match="psrc and n must be the same size"):
        pe.StokesDLP_ewald_2p(x, x, x, x, np.zeros(11), np.zeros(11), x, x, 1.0, 1.0)
    ps = np.zeros((10, 2))
    with pytest.raises(ValueError, match="psrc and f must be the same size"):
        pe.autotune("slp", ps, ps, np.zeros((8, 2)), None, 1.0, 1.0, tol=1e-8)
    with pytest.raises(ValueError, match="psrc and n must be the same size"):
        pe.autotune("dlp", ps, ps, np.zeros((10, 2)), np.zeros((3, 2)), 1.0, 1.0, tol=1e-8)


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_fixed_grid_mode_inverts_the_driver_estimates(kind):
    """Fixed-grid parameters (grid given, xi and cells derived): the drivers'
    own Newton solves, run forward on the result, give back the grid (k_inf =
    M/2) and a cutoff no wider than the cells; one more cell per side would
    be narrower than the cutoff the estimate needs.  The C selection agrees
    with the Python restatement."""
    lib = _capi.load()
    find_xi = Pm.slp_find_xi if kind == "slp" else Pm.dlp_find_xi
    find_k = Pm.slp_find_kinfb if kind == "slp" else Pm.dlp_find_kinfb
    est_r = Pm.slp_real_estimate if kind == "slp" else Pm.dlp_real_estimate
    est_k = Pm.slp_kspace_estimate if kind == "slp" else Pm.dlp_kspace_estimate
    for (N, Lx, Ly, tol, M, seed) in ((4000, 1.0, 1.0, 1e-12, 128, 1), (20000, 2.0, 1.0, 1e-10, 256, 2),
                                      (1000, 1.0, 1.0, 1e-8, 64, 3)):
        rng = np.random.default_rng(seed)
        f = rng.standard_normal((N, 2))
        Q = float(np.sum(f * f)) + 1
        p = Pm.select_params_grid(kind, N, N, f, Lx, Ly, M=M, tol=tol)
        A, B = Pm.matlab_rat(Lx / Ly)
        assert (p.Mx, p.My) == (A * M, B * M)
        # the k-space estimate (Lx passed twice, as the drivers do) is tol at
        # k_inf = M/2, and the drivers' Newton solve lands on that grid (it
        # stops anywhere the estimate is within [0, 2 tol], so one below)
        k = 2 * np.pi * (M // 2) / Lx
        assert abs(est_k(Q, Lx, Lx, p.xi, k) / tol - 1) < 1e-9
        assert find_k(Q, Lx, Lx, p.xi, tol) in (M // 2 - 1, M // 2)
        # the real-space estimate is at most tol at the cell side, and one more
        # cell per side would make the cells narrower than the cutoff it needs
        assert est_r(Q, Lx, Ly, p.xi, p.rc) <= tol * (1 + 1e-9)
        a = p.nside_x // A
        assert est_r(Q, Lx, Ly, p.xi, Lx / ((a + 1) * A)) > tol
        assert find_xi(Q, Lx, Ly, p.rc, tol) <= p.xi * (1 + 1e-9)
        prm = _capi.Se2pParams()
        prm.tol, prm.M, prm.fixed_grid = tol, M, 1
        assert lib.se2p_select_params(0 if kind == "slp" else 1, N, N, Q, Lx, Ly, ctypes.byref(prm)) == 0
        assert (prm.nside_x, prm.nside_y, prm.Mx, prm.My, prm.kinf) == \
            (p.nside_x, p.nside_y, p.Mx, p.My, p.kinf)
        assert abs(prm.xi - p.xi) <= 1e-12 * p.xi and abs(prm.eta - p.eta) <= 1e-12 * p.eta
        assert prm.converged == 1
    prm = _capi.Se2pParams()
    prm.tol, prm.M, prm.fixed_grid = 1e-10, 63, 1
    assert lib.se2p_select_params(0, 100, 100, 200.0, 1.0, 1.0, ctypes.byref(prm)) == _capi.SE2P_ERR_GRID


def test_fixed_grid_config5_values():
    """N = 4e6 uniform points, tol 1e-12, grid 2048^2: the grid is below the
    drivers' 2 k_inf (2440 and 3060), so xi drops and the cells grow."""
    f = np.random.default_rng(5).standard_normal((4_000_000, 2))
    s = Pm.select_params_grid("slp", 4_000_000, 4_000_000, f, 1.0, 1.0, M=2048, tol=1e-12)
    d = Pm.select_params_grid("dlp", 4_000_000, 4_000_000, f, 1.0, 1.0, M=2048, tol=1e-12)
    assert (s.nside_x, s.Mx) == (99, 2048) and abs(s.xi - 564.65) < 0.01 and s.eta < 1
    assert (d.nside_x, d.Mx) == (80, 2048) and abs(d.xi - 498.92) < 0.01 and d.eta < 1
