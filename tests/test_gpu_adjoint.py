"""Backward entry points (adjoints) on the GPU vs the oracle's adjoints.

Bar: max-norm relative error <= 1e-12 per output array (dL/dpsrc, dL/dptar,
dL/df, dL/dn) against oracle/se2p_oracle.c or_*_adj on the same seeded
inputs.  At full size (stresslet on 200 ellipses, N=1e5, 512^2) du is
nonzero on a subset of targets so that the oracle finishes in seconds, and
size-independent properties are checked on the full problem: the adjoint
identity <dL/df, v> = <du, u(v)> with the GPU's own forward sums, and the
translation invariance of the real-space gradients.
"""
import numpy as np
import pytest
import torch

from oracle import cref
from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W
from tests.conftest import relerr

TOL = 1e-12

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _np(d):
    return {k: v.cpu().numpy() for k, v in d.items()}


def _say(msg):
    print(msg, flush=True)


def _gpu_adj(pe, kind, ps, pt, f, nv, du, p, Lx, Ly, same=False, wrt=("psrc", "ptar", "f", "n")):
    dps = _dev(ps)
    dpt = dps if same else _dev(pt)
    if kind == "slp_real":
        g = pe.mex_stokes_slp_real_backward(dps, dpt, _dev(f), p.xi, p.nside_x, p.nside_y, Lx, Ly,
                                            _dev(du), wrt=wrt)
    elif kind == "dlp_real":
        g = pe.mex_stokes_dlp_real_backward(dps, dpt, _dev(f), _dev(nv), p.xi, p.nside_x,
                                            p.nside_y, Lx, Ly, _dev(du), wrt=wrt)
    elif kind == "slp_kspace":
        g = pe.mex_stokes_slp_kspace_backward(dps, dpt, p.xi, p.eta, _dev(f), p.Mx, p.My, Lx, Ly,
                                              p.w, p.P, _dev(du), wrt=wrt)
    else:
        g = pe.mex_stokes_dlp_kspace_backward(dps, dpt, p.xi, p.eta, _dev(f), _dev(nv), p.Mx,
                                              p.My, Lx, Ly, p.w, p.P, _dev(du), wrt=wrt)
    return _np(g)


def _cpu_adj(kind, ps, pt, f, nv, du, p, Lx, Ly):
    if kind == "slp_real":
        g = cref.slp_real_adj(ps, pt, f, p.xi, p.nside_x, p.nside_y, Lx, Ly, du)
    elif kind == "dlp_real":
        g = cref.dlp_real_adj(ps, pt, f, nv, p.xi, p.nside_x, p.nside_y, Lx, Ly, du)
    elif kind == "slp_kspace":
        g = cref.slp_kspace_adj(ps, pt, p.xi, p.eta, f, p.Mx, p.My, Lx, Ly, p.w, p.P, du)
    else:
        g = cref.dlp_kspace_adj(ps, pt, p.xi, p.eta, f, nv, p.Mx, p.My, Lx, Ly, p.w, p.P, du)
    names = {"dpsrc": "psrc", "dptar": "ptar", "df": "f", "dn": "n"}
    return {names[k]: v for k, v in g.items()}


KINDS = ["slp_real", "slp_kspace", "dlp_real", "dlp_kspace"]


def _check_all(pe, ps, pt, f, nv, du, pr, pd, Lx, Ly, label, same=False):
    worst = {}
    for kind in KINDS:
        p = pr if kind.startswith("slp") else pd
        _say(f"[{label}] {kind} Ns={len(ps)} Nt={len(pt)}")
        g = _gpu_adj(pe, kind, ps, pt, f, nv, du, p, Lx, Ly, same=same)
        c = _cpu_adj(kind, ps, pt, f, nv, du, p, Lx, Ly)
        assert set(g) == set(c), (kind, set(g), set(c))
        for k in c:
            e = relerr(g[k], c[k])
            worst[(kind, k)] = e
            assert e <= TOL, (label, kind, k, e)
    _say(f"[{label}] worst {max(worst.values()):.2e}")


def test_adjoint_unit_box_separate_targets(gpu):
    w = W.uniform_box(1000, seed=1)
    rng = np.random.default_rng(21)
    pt = Pm.wrap(rng.random((600, 2)), 1.0)
    du = rng.standard_normal((600, 2))
    pr = Pm.select_params("slp", 1000, 600, w.f, 1.0, 1.0, tol=1e-10, M=64)
    pd = Pm.select_params("dlp", 1000, 600, w.f, 1.0, 1.0, tol=1e-10, M=64)
    _check_all(gpu, w.psrc, pt, w.f, w.n, du, pr, pd, 1.0, 1.0, "box")


def test_adjoint_targets_are_sources(gpu):
    # one array for both: shared cell lists and bins, coincident self pairs
    w = W.uniform_box(800, seed=4)
    du = np.random.default_rng(22).standard_normal((800, 2))
    pr = Pm.select_params("slp", 800, 800, w.f, 1.0, 1.0, tol=1e-10, M=48)
    pd = Pm.select_params("dlp", 800, 800, w.f, 1.0, 1.0, tol=1e-10, M=48)
    _check_all(gpu, w.psrc, w.psrc, w.f, w.n, du, pr, pd, 1.0, 1.0, "same", same=True)


def test_adjoint_rectangular_box(gpu):
    rng = np.random.default_rng(23)
    Lx, Ly = 1.0, 2.0
    Ns, Nt = 500, 300
    ps = np.stack([Pm.wrap(Lx * rng.random(Ns), Lx), Pm.wrap(Ly * rng.random(Ns), Ly)], 1)
    pt = np.stack([Pm.wrap(Lx * rng.random(Nt), Lx), Pm.wrap(Ly * rng.random(Nt), Ly)], 1)
    f = rng.standard_normal((Ns, 2))
    nv = rng.standard_normal((Ns, 2))
    nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    du = rng.standard_normal((Nt, 2))
    pr = Pm.select_params("slp", Ns, Nt, f, Lx, Ly, tol=1e-12)
    pd = Pm.select_params("dlp", Ns, Nt, f, Lx, Ly, tol=1e-12)
    _check_all(gpu, ps, pt, f, nv, du, pr, pd, Lx, Ly, "rect")


def test_adjoint_small_nside_and_small_grids(gpu):
    # nside 1 and 2 revisit a cell under several image offsets; grids 24 and
    # 14 make the Gaussian support wrap onto itself
    rng = np.random.default_rng(24)
    P, L = 24, 1.0
    m = 0.95 * np.sqrt(np.pi * P)
    for N, nside, MM in ((7, 1, 24), (40, 2, 14), (120, 2, 32)):
        ps = Pm.wrap(rng.random((N, 2)), L)
        pt = Pm.wrap(rng.random((N + 3, 2)), L)
        f = rng.standard_normal((N, 2))
        nv = rng.standard_normal((N, 2))
        nv /= np.linalg.norm(nv, axis=1, keepdims=True)
        du = rng.standard_normal((N + 3, 2))
        w = P * L / MM / 2
        xi = np.sqrt(0.5) * m / (2 * w)
        p = Pm.EwaldParams(P=P, Nb=0.0, tol=0.0, nside_x=nside, nside_y=nside, rc=L / nside, xi=xi,
                           kinf=0, Mx=MM, My=MM, w=w, eta=0.5, A=1, B=1)
        _check_all(gpu, ps, pt, f, nv, du, p, p, L, L, f"N{N}-ns{nside}-M{MM}")


def test_adjoint_partial_outputs_and_empty_sides(gpu):
    w = W.uniform_box(300, seed=6)
    rng = np.random.default_rng(25)
    du = rng.standard_normal((300, 2))
    pd = Pm.select_params("dlp", 300, 300, w.f, 1.0, 1.0, tol=1e-10, M=32)
    full = _gpu_adj(gpu, "dlp_kspace", w.psrc, w.ptar, w.f, w.n, du, pd, 1.0, 1.0)
    for wrt in (("ptar",), ("f",), ("psrc", "n")):
        part = _gpu_adj(gpu, "dlp_kspace", w.psrc, w.ptar, w.f, w.n, du, pd, 1.0, 1.0, wrt=wrt)
        assert set(part) == set(wrt)
        for k in wrt:
            assert np.array_equal(part[k], full[k]), (wrt, k)
    full = _gpu_adj(gpu, "slp_real", w.psrc, w.ptar, w.f, w.n, du, pd, 1.0, 1.0)
    part = _gpu_adj(gpu, "slp_real", w.psrc, w.ptar, w.f, w.n, du, pd, 1.0, 1.0, wrt=("f",))
    assert np.array_equal(part["f"], full["f"])
    e = np.zeros((0, 2))
    for kind in KINDS:
        g = _gpu_adj(gpu, kind, e, w.ptar, e, e, du, pd, 1.0, 1.0)
        assert np.all(g["ptar"] == 0) and g["psrc"].shape == (0, 2), kind
        g = _gpu_adj(gpu, kind, w.psrc, e, w.f, w.n, e, pd, 1.0, 1.0)
        assert np.all(g["psrc"] == 0) and np.all(g["f"] == 0), kind


def test_adjoint_ellipses_full_size(gpu):
    """Stresslet on 200 ellipses x 500 nodes, grid 512^2: du nonzero on 2000
    targets, every gradient compared with the oracle; then the full problem's
    adjoint identity and real-space translation invariance."""
    w, p = W.dlp_ellipses_config()
    rng = np.random.default_rng(56)
    n = len(w.psrc)
    du = np.zeros((n, 2))
    sub = rng.choice(n, 2000, replace=False)
    du[sub] = rng.standard_normal((2000, 2))
    for kind in ("dlp_real", "dlp_kspace"):
        _say(f"[ellipses] {kind} sparse du")
        g = _gpu_adj(gpu, kind, w.psrc, w.psrc, w.f, w.n, du, p, 1.0, 1.0, same=True)
        c = _cpu_adj(kind, w.psrc, w.psrc, w.f, w.n, du, p, 1.0, 1.0)
        for k in c:
            e = relerr(g[k], c[k])
            _say(f"[ellipses] {kind} {k} {e:.2e}")
            assert e <= TOL, (kind, k, e)

    # full du: properties that hold at any size
    du = rng.standard_normal((n, 2))
    v = rng.standard_normal((n, 2))
    ps, f, nv = _dev(w.psrc), _dev(w.f), _dev(w.n)
    for kind in ("dlp_real", "dlp_kspace"):
        _say(f"[ellipses] {kind} full du: adjoint identity")
        g = _gpu_adj(gpu, kind, w.psrc, w.psrc, w.f, w.n, du, p, 1.0, 1.0, same=True)
        if kind == "dlp_real":
            uf = gpu.mex_stokes_dlp_real(ps, ps, _dev(v), nv, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
            un = gpu.mex_stokes_dlp_real(ps, ps, f, _dev(v), p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
        else:
            uf = gpu.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, _dev(v), nv, p.Mx, p.My, 1.0, 1.0, p.w, p.P)
            un = gpu.mex_stokes_dlp_kspace(ps, ps, p.xi, p.eta, f, _dev(v), p.Mx, p.My, 1.0, 1.0, p.w, p.P)
        lf = float(np.sum(du * uf.cpu().numpy()))
        ln = float(np.sum(du * un.cpu().numpy()))
        assert abs(np.sum(g["f"] * v) - lf) <= 1e-12 * np.sum(np.abs(g["f"] * v)), (kind, lf)
        assert abs(np.sum(g["n"] * v) - ln) <= 1e-12 * np.sum(np.abs(g["n"] * v)), (kind, ln)
        if kind == "dlp_real":
            tot = g["psrc"].sum(axis=0) + g["ptar"].sum(axis=0)
            assert np.abs(tot).max() <= 1e-12 * np.abs(g["psrc"]).sum(), tot


def test_adjoint_one_pass_equals_two_passes(gpu):
    """Targets = sources with both sides wanted: the Fourier-space backward
    runs both passes as one (densities and du spread together, one transform
    pair with both filters, one interpolation).  Against the two-pass form
    (se2p_set_adjoint_passes(ctx, 2)): within 1e-13 on every output, and both
    within 1e-12 of the oracle; grids of 48 (strip spreading, tile
    interpolation) and 14 and 24 (support wrapping onto itself, generic
    spreading); a partial output set gives the full call's values bitwise."""
    pe = gpu
    rng = np.random.default_rng(57)
    P, L = 24, 1.0
    m = 0.95 * np.sqrt(np.pi * P)
    one, two = pe.Context(0), pe.Context(0)
    two.set_adjoint_passes(2)
    for N, MM in ((800, 48), (40, 14), (150, 24)):
        x = Pm.wrap(rng.random((N, 2)), L)
        f = rng.standard_normal((N, 2))
        nv = rng.standard_normal((N, 2))
        nv /= np.linalg.norm(nv, axis=1, keepdims=True)
        du = rng.standard_normal((N, 2))
        w = P * L / MM / 2
        xi = np.sqrt(0.5) * m / (2 * w)
        ps, ff, nn, dd = _dev(x), _dev(f), _dev(nv), _dev(du)
        for kind in ("slp_kspace", "dlp_kspace"):
            _say(f"[one pass] {kind} N={N} M={MM}")

            def run(ctx, wrt=("psrc", "ptar", "f", "n")):
                if kind == "slp_kspace":
                    g = pe.mex_stokes_slp_kspace_backward(ps, ps, xi, 0.5, ff, MM, MM, L, L, w, P, dd,
                                                          wrt=tuple(k for k in wrt if k != "n"),
                                                          ctx=ctx)
                else:
                    g = pe.mex_stokes_dlp_kspace_backward(ps, ps, xi, 0.5, ff, nn, MM, MM, L, L, w, P,
                                                          dd, wrt=wrt, ctx=ctx)
                return _np(g)

            g1, g2 = run(one), run(two)
            if kind == "slp_kspace":
                c = cref.slp_kspace_adj(x, x, xi, 0.5, f, MM, MM, L, L, w, P, du)
            else:
                c = cref.dlp_kspace_adj(x, x, xi, 0.5, f, nv, MM, MM, L, L, w, P, du)
            names = {"dpsrc": "psrc", "dptar": "ptar", "df": "f", "dn": "n"}
            for k, v in c.items():
                k = names[k]
                e12, e1 = relerr(g1[k], g2[k]), relerr(g1[k], v)
                assert e12 <= 1e-13 and e1 <= TOL and relerr(g2[k], v) <= TOL, (kind, N, MM, k, e12, e1)
            part = run(one, wrt=("ptar", "f"))
            for k in part:
                assert np.array_equal(part[k], g1[k]), (kind, k)
