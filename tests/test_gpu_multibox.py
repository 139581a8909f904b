"""Multi-box entry points on the GPU.

Every box of a multi-box call must give exactly (bitwise) what the
single-box entry point gives on that box: the real-space sums of all boxes
share one sort and one pair-kernel launch, and the stable sort keeps each
box's summation order.  A small mixed set is also compared with the oracle
at 1e-12, and the full many-box case (32 boxes, N from 1e3 to 2e5) with the
single-box calls.
"""
import numpy as np
import pytest
import torch

from oracle import cref
from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W
from tests.conftest import relerr

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _say(msg):
    print(msg, flush=True)


def _mixed_boxes():
    """Six boxes: different sizes, a rectangular box, separate targets, a box
    without targets and one without sources."""
    rng = np.random.default_rng(40)
    boxes = []
    for (Lx, Ly, N, Nt, M) in ((1.0, 1.0, 900, None, 64), (2.0, 1.0, 600, 250, 0), (0.5, 0.5, 300, None, 0),
                               (3.0, 3.0, 1200, 400, 0), (1.0, 1.0, 50, 0, 32), (1.0, 1.0, 0, 20, 32)):
        x = np.stack([Pm.wrap(Lx * rng.random(N), Lx), Pm.wrap(Ly * rng.random(N), Ly)], 1)
        t = x if Nt is None else np.stack([Pm.wrap(Lx * rng.random(Nt), Lx), Pm.wrap(Ly * rng.random(Nt), Ly)], 1)
        f = rng.standard_normal((N, 2))
        nv = rng.standard_normal((N, 2))
        nv /= np.maximum(np.linalg.norm(nv, axis=1, keepdims=True), 1e-300)
        p = Pm.select_params("slp", max(N, 1), max(len(t), 1), f if N else np.zeros((1, 2)), Lx, Ly,
                             tol=1e-10, M=M or None)
        boxes.append({"psrc": x, "ptar": t, "f": f, "n": nv, "Lx": Lx, "Ly": Ly, "xi": p.xi,
                      "nside_x": p.nside_x, "nside_y": p.nside_y, "eta": p.eta, "w": p.w,
                      "Mx": p.Mx, "My": p.My, "P": p.P})
    return boxes


def _single(pe, kind, b):
    ps, pt, f, n = _dev(b["psrc"]), _dev(b["ptar"]), _dev(b["f"]), _dev(b["n"])
    if kind == "slp_real":
        return pe.mex_stokes_slp_real(ps, pt, f, b["xi"], b["nside_x"], b["nside_y"], b["Lx"], b["Ly"])
    if kind == "dlp_real":
        return pe.mex_stokes_dlp_real(ps, pt, f, n, b["xi"], b["nside_x"], b["nside_y"], b["Lx"], b["Ly"])
    if kind == "slp_kspace":
        return pe.mex_stokes_slp_kspace(ps, pt, b["xi"], b["eta"], f, b["Mx"], b["My"], b["Lx"], b["Ly"],
                                        b["w"], b["P"])
    return pe.mex_stokes_dlp_kspace(ps, pt, b["xi"], b["eta"], f, n, b["Mx"], b["My"], b["Lx"], b["Ly"],
                                    b["w"], b["P"])


def _oracle(kind, b):
    if kind == "slp_real":
        return cref.slp_real(b["psrc"], b["ptar"], b["f"], b["xi"], b["nside_x"], b["nside_y"], b["Lx"], b["Ly"])
    if kind == "dlp_real":
        return cref.dlp_real(b["psrc"], b["ptar"], b["f"], b["n"], b["xi"], b["nside_x"], b["nside_y"],
                             b["Lx"], b["Ly"])
    if kind == "slp_kspace":
        return cref.slp_kspace(b["psrc"], b["ptar"], b["xi"], b["eta"], b["f"], b["Mx"], b["My"], b["Lx"],
                               b["Ly"], b["w"], b["P"])
    return cref.dlp_kspace(b["psrc"], b["ptar"], b["xi"], b["eta"], b["f"], b["n"], b["Mx"], b["My"],
                           b["Lx"], b["Ly"], b["w"], b["P"])


def _multi(pe, kind, boxes):
    dev_boxes = []
    for b in boxes:
        d = dict(b)
        for k in ("psrc", "f", "n"):
            d[k] = _dev(b[k])
        d["ptar"] = d["psrc"] if b["ptar"] is b["psrc"] else _dev(b["ptar"])
        if kind.startswith("slp"):
            d.pop("n")
        dev_boxes.append(d)
    fn = {"slp_real": pe.stokes_slp_real_multibox, "dlp_real": pe.stokes_dlp_real_multibox,
          "slp_kspace": pe.stokes_slp_kspace_multibox, "dlp_kspace": pe.stokes_dlp_kspace_multibox}[kind]
    return [u.cpu().numpy() for u in fn(dev_boxes)]


KINDS = ["slp_real", "dlp_real", "slp_kspace", "dlp_kspace"]


@pytest.mark.parametrize("kind", KINDS)
def test_multibox_equals_single_and_oracle(gpu, kind):
    boxes = _mixed_boxes()
    _say(f"[multibox mixed] {kind} nbox={len(boxes)}")
    outs = _multi(gpu, kind, boxes)
    for i, b in enumerate(boxes):
        assert outs[i].shape == (len(b["ptar"]), 2)
        one = _single(gpu, kind, b).cpu().numpy()
        assert np.array_equal(outs[i], one), (kind, i, relerr(outs[i], one))
        if len(b["ptar"]) and len(b["psrc"]):
            e = relerr(outs[i], _oracle(kind, b))
            assert e <= 1e-12, (kind, i, e)
        else:
            assert np.all(outs[i] == 0)


@pytest.mark.parametrize("kernel", ["slp", "dlp"])
def test_multibox_many_boxes_equals_single(gpu, kernel):
    """The 32-box workload (L ~ U[0.5, 4], N log-uniform in [1e3, 2e5]), each
    kernel with its own driver parameters per box."""
    boxes = W.multibox_config(seed=4, nbox=32, kind=kernel)
    _say(f"[multibox 32 {kernel}] N total={sum(len(b['psrc']) for b in boxes)}")
    for kind in (kernel + "_real", kernel + "_kspace"):
        outs = _multi(gpu, kind, boxes)
        for i, b in enumerate(boxes):
            one = _single(gpu, kind, b).cpu().numpy()
            assert np.array_equal(outs[i], one), (kind, i, relerr(outs[i], one))
        _say(f"[multibox 32] {kind} identical to single-box calls")


@pytest.mark.parametrize("targets", ["mixed", "sources"])
def test_multibox_packs_with_mixed_grids(gpu, targets):
    """20 boxes, more than one packed launch (16 boxes each): grids from 32
    to 320 points a side, so that some boxes take the packed strip spreading
    and others the single-box kernels (grids narrower than P + 16), square
    and rectangular; targets = sources in every other box ("mixed": each box
    sorted on its own) or in every box ("sources": one sort over all boxes).
    Every box bitwise equal to its single-box call, Fourier space of both
    kernels."""
    rng = np.random.default_rng(46)
    boxes = []
    for i in range(20):
        Lx = float(rng.choice([0.5, 1.0, 2.0]))
        Ly = Lx if i % 3 else 2 * Lx
        N = int(rng.integers(50, 2000))
        x = np.stack([Pm.wrap(Lx * rng.random(N), Lx), Pm.wrap(Ly * rng.random(N), Ly)], 1)
        t = x if (i % 2 or targets == "sources") else np.stack(
            [Pm.wrap(Lx * rng.random(N // 2 + 1), Lx), Pm.wrap(Ly * rng.random(N // 2 + 1), Ly)], 1)
        f = rng.standard_normal((N, 2))
        nv = rng.standard_normal((N, 2))
        nv /= np.linalg.norm(nv, axis=1, keepdims=True)
        M = int(rng.choice([20, 32, 48, 64, 96, 160]))
        p = Pm.select_params("dlp", N, len(t), f, Lx, Ly, tol=1e-10, M=M)
        if p.eta >= 1:   # grid below the driver's: take the driver's
            p = Pm.select_params("dlp", N, len(t), f, Lx, Ly, tol=1e-10)
        boxes.append({"psrc": x, "ptar": t, "f": f, "n": nv, "Lx": Lx, "Ly": Ly, "xi": p.xi,
                      "nside_x": p.nside_x, "nside_y": p.nside_y, "eta": p.eta, "w": p.w,
                      "Mx": p.Mx, "My": p.My, "P": p.P})
    narrow = sum(min(b["Mx"], b["My"]) < b["P"] + 16 for b in boxes)
    _say(f"[multibox packs {targets}] 20 boxes, {narrow} with grids narrower than P + 16")
    assert 0 < narrow < 20
    for kind in ("slp_kspace", "dlp_kspace"):
        outs = _multi(gpu, kind, boxes)
        for i, b in enumerate(boxes):
            one = _single(gpu, kind, b).cpu().numpy()
            assert np.array_equal(outs[i], one), (kind, i, relerr(outs[i], one))
        _say(f"[multibox packs] {kind} identical to single-box calls")


def test_multibox_autotune(gpu):
    """se2p_autotune_multibox on 8 boxes of the many-box workload (Stokeslet,
    tol 1e-10): the parameters it writes give a total (real + Fourier) within
    5e-10 of the drivers' parameters' total on every box (the Ewald total
    does not depend on the split; each meets tol), and the smallest box's
    real- and Fourier-space sums at the tuned parameters match the oracle at
    1e-12."""
    pe = gpu
    boxes = W.multibox_config(seed=4, nbox=8)
    _say(f"[multibox autotune] N total={sum(len(b['psrc']) for b in boxes)}")
    dev_boxes = []
    for b in boxes:
        d = {k: b[k] for k in ("Lx", "Ly")}
        d["psrc"] = d["ptar"] = _dev(b["psrc"])
        d["f"] = _dev(b["f"])
        dev_boxes.append(d)
    ctx = pe.Context(0)
    info = pe.autotune_multibox("slp", dev_boxes, tol=1e-10, ctx=ctx)
    _say(f"  drivers {info['drivers_ms']:.3f} ms, tuned {info['tuned_ms']:.3f} ms, split "
         f"{info['spread_chunk']}, kept drivers {info['kept_drivers']}")
    assert info["spread_chunk"] in (0, 128, 512)
    assert info["tuned_ms"] <= info["drivers_ms"]
    tuned = []
    for b, d in zip(boxes, dev_boxes):
        t = dict(b)
        t.update({k: d[k] for k in ("xi", "nside_x", "nside_y", "eta", "w", "Mx", "My", "P")})
        tuned.append(t)
    ur_t, uk_t = _multi(pe, "slp_real", tuned), _multi(pe, "slp_kspace", tuned)
    ur_d, uk_d = _multi(pe, "slp_real", boxes), _multi(pe, "slp_kspace", boxes)
    for i in range(len(boxes)):
        e = relerr(ur_t[i] + uk_t[i], ur_d[i] + uk_d[i])
        assert e <= 5e-10, (i, e, tuned[i]["nside_x"], boxes[i]["nside_x"])
    i = int(np.argmin([len(b["psrc"]) for b in boxes]))
    assert relerr(ur_t[i], _oracle("slp_real", tuned[i])) <= 1e-12
    assert relerr(uk_t[i], _oracle("slp_kspace", tuned[i])) <= 1e-12
