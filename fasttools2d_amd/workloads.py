"""Synthetic workloads for parity tests and benchmarks (seeded numpy).

* ``uniform_box``      -- N uniform points in [0,L)^2 wrapped to [-L/2,L/2),
                          N(0,1) densities, unit normals at uniform angles
                          (the Stokeslet plain case: N=1000, unit box).
* ``ellipses``         -- non-overlapping ellipses with trapezoid-weighted
                          smooth densities (the stresslet case: 200 ellipses x
                          500 nodes, semi-axes U[0.01,0.03]).
* ``stresslet_circle`` -- the stresslet-identity geometry of
                          PeriodicEwald/tests/stresslet_identity_test.m:31-46.

Every point set is returned in the reference's (N, 2) = MATLAB 2xN layout.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import params as _params


@dataclass
class Workload:
    name: str
    psrc: np.ndarray          # (Ns, 2), wrapped into [-L/2, L/2)
    ptar: np.ndarray          # (Nt, 2)
    f: np.ndarray             # (Ns, 2)
    n: np.ndarray             # (Ns, 2) unit normals
    Lx: float
    Ly: float
    psrc_unwrapped: np.ndarray | None = None
    meta: dict = field(default_factory=dict)


def uniform_box(N: int, seed: int, L: float = 1.0, targets_equal_sources: bool = True,
                Nt: int | None = None) -> Workload:
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, L, size=(N, 2))
    f = rng.standard_normal((N, 2))
    ang = rng.uniform(0.0, 2 * np.pi, size=N)
    nv = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    ps = np.stack([_params.wrap(x[:, 0], L), _params.wrap(x[:, 1], L)], axis=1)
    if targets_equal_sources:
        pt = ps.copy()
    else:
        xt = rng.uniform(0.0, L, size=(Nt or N, 2))
        pt = np.stack([_params.wrap(xt[:, 0], L), _params.wrap(xt[:, 1], L)], axis=1)
    return Workload(f"uniform_N{N}_seed{seed}", ps, pt, f, nv, L, L, psrc_unwrapped=x)


def ellipses(n_ell: int = 200, n_nodes: int = 500, seed: int = 2, L: float = 1.0,
             amin: float = 0.01, amax: float = 0.03, gap: float = 0.005,
             kmax: int = 8) -> Workload:
    """Non-overlapping ellipses (bounding circles at least ``gap`` apart,
    periodic), ``n_nodes`` equispaced parameter nodes each, outward unit
    normals, density ``sum_{k<=kmax} a_k cos k t + b_k sin k t`` per component
    with ``a_k, b_k ~ N(0,1)/(1+k^2)``, multiplied by the trapezoid weight
    ``|x'(t)| 2 pi / n_nodes``."""
    rng = np.random.default_rng(seed)
    ab = rng.uniform(amin, amax, size=(n_ell, 2))
    phi = rng.uniform(0.0, np.pi, size=n_ell)
    R = ab.max(axis=1)
    order = np.argsort(-R)              # place the large ones first
    centres = np.zeros((n_ell, 2))
    placed = []
    for e in order:
        for _ in range(200000):
            c = rng.uniform(0.0, L, size=2)
            if placed:
                pc = centres[placed]
                d = c[None, :] - pc
                d -= L * np.round(d / L)
                dist = np.sqrt((d * d).sum(axis=1))
                if np.any(dist < R[placed] + R[e] + gap):
                    continue
            centres[e] = c
            placed.append(e)
            break
        else:
            raise RuntimeError("could not place all ellipses")
    t = 2 * np.pi * np.arange(n_nodes) / n_nodes
    xs, nvs, fs = [], [], []
    ks = np.arange(kmax + 1)
    for e in range(n_ell):
        a, b = ab[e]
        cph, sph = math.cos(phi[e]), math.sin(phi[e])
        lx, ly = a * np.cos(t), b * np.sin(t)
        x = centres[e, 0] + cph * lx - sph * ly
        y = centres[e, 1] + sph * lx + cph * ly
        # outward normal  ~ R(phi) (b cos t, a sin t)
        nx0, ny0 = b * np.cos(t), a * np.sin(t)
        nn = np.sqrt(nx0 * nx0 + ny0 * ny0)
        nx, ny = (cph * nx0 - sph * ny0) / nn, (sph * nx0 + cph * ny0) / nn
        speed = np.sqrt((a * np.sin(t)) ** 2 + (b * np.cos(t)) ** 2)
        wgt = speed * 2 * np.pi / n_nodes
        dens = []
        for _comp in range(2):
            ca = rng.standard_normal(kmax + 1) / (1 + ks ** 2)
            cb = rng.standard_normal(kmax + 1) / (1 + ks ** 2)
            dens.append((ca[None, :] * np.cos(np.outer(t, ks)) +
                         cb[None, :] * np.sin(np.outer(t, ks))).sum(axis=1))
        xs.append(np.stack([x, y], axis=1))
        nvs.append(np.stack([nx, ny], axis=1))
        fs.append(np.stack([dens[0] * wgt, dens[1] * wgt], axis=1))
    xu = np.concatenate(xs)
    nv = np.concatenate(nvs)
    f = np.concatenate(fs)
    ps = np.stack([_params.wrap(xu[:, 0], L), _params.wrap(xu[:, 1], L)], axis=1)
    return Workload(f"ellipses_{n_ell}x{n_nodes}_seed{seed}", ps, ps.copy(), f, nv, L, L,
                    psrc_unwrapped=xu, meta={"centres": centres, "semi_axes": ab, "phi": phi})


def stresslet_circle(nsrc: int = 256, ntar: int = 16, r: float = 1.0, comp: int = 0) -> Workload:
    """Circle of radius r in a 4r x 4r box, density h*e_comp, 16 targets at
    r/2 and 16 at 3r/2 (stresslet_identity_test.m:31-46)."""
    L = 4 * r
    h = 2 * np.pi / nsrc
    th = h * np.arange(1, nsrc + 1)
    xs = np.stack([r * np.cos(th), r * np.sin(th)], axis=1)
    nv = np.stack([np.cos(th), np.sin(th)], axis=1)
    f = np.zeros((nsrc, 2))
    f[:, comp] = h
    ht = 2 * np.pi / ntar
    tt = ht * np.arange(1, ntar + 1)
    pin = np.stack([0.5 * r * np.cos(tt), 0.5 * r * np.sin(tt)], axis=1)
    pout = np.stack([1.5 * r * np.cos(tt), 1.5 * r * np.sin(tt)], axis=1)
    pt = np.concatenate([pin, pout])
    ps = np.stack([_params.wrap(xs[:, 0], L), _params.wrap(xs[:, 1], L)], axis=1)
    ptw = np.stack([_params.wrap(pt[:, 0], L), _params.wrap(pt[:, 1], L)], axis=1)
    return Workload(f"stresslet_circle_{nsrc}", ps, ptw, f, nv, L, L, psrc_unwrapped=xs)


def slp_plain_config(seed: int = 1, N: int = 1000, tol: float = 1e-10, M: int = 64):
    """Plain Stokeslet: N uniform points, targets = sources, grid M^2."""
    w = uniform_box(N, seed)
    p = _params.select_params("slp", N, N, w.f, w.Lx, w.Ly, tol=tol, M=M)
    return w, p


def dlp_plain_config(seed: int = 1, N: int = 1000, tol: float = 1e-10, M: int = 64):
    w = uniform_box(N, seed)
    p = _params.select_params("dlp", N, N, w.f, w.Lx, w.Ly, tol=tol, M=M)
    return w, p


def dlp_ellipses_config(seed: int = 2, n_ell: int = 200, n_nodes: int = 500,
                        tol: float = 1e-10, M: int = 512):
    """Stresslet on 200 ellipses x 500 nodes (N = 1e5), grid 512^2, tol 1e-10."""
    w = ellipses(n_ell, n_nodes, seed)
    N = len(w.psrc)
    p = _params.select_params("dlp", N, N, w.f, w.Lx, w.Ly, tol=tol, M=M)
    return w, p


def multibox_config(seed: int = 4, nbox: int = 32, L_range=(0.5, 4.0), N_range=(1e3, 2e5),
                    tol: float = 1e-10, kind: str = "slp", P: int = 24):
    """Many independent periodic boxes: square boxes with side L ~ U[0.5, 4],
    N = round(exp(U[ln 1e3, ln 2e5])) uniform points each (targets = sources),
    N(0,1) densities, unit normals at uniform angles, every box's Ewald
    parameters from the driver formulas at tol.  Returns a list of box dicts
    for the multi-box entry points."""
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(nbox):
        L = float(rng.uniform(*L_range))
        N = int(round(math.exp(rng.uniform(math.log(N_range[0]), math.log(N_range[1])))))
        x = _params.wrap(rng.uniform(0.0, L, size=(N, 2)), L)
        f = rng.standard_normal((N, 2))
        th = rng.uniform(0.0, 2 * math.pi, N)
        nv = np.stack([np.cos(th), np.sin(th)], axis=1)
        p = _params.select_params(kind, N, N, f, L, L, P=P, tol=tol)
        boxes.append({"psrc": x, "ptar": x, "f": f, "n": nv, "Lx": L, "Ly": L, "xi": p.xi,
                      "nside_x": p.nside_x, "nside_y": p.nside_y, "eta": p.eta, "w": p.w,
                      "Mx": p.Mx, "My": p.My, "P": p.P})
    return boxes
