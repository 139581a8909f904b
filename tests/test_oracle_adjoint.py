"""Pin the oracle's adjoints (oracle/se2p_oracle.c, or_*_adj) before trusting them.

The reference has no adjoint, so there are no golden vectors; the pins are
what the mathematics fixes:
* the density gradients through the adjoint identity: each sum is linear in
  f (and, for the stresslet, in n), so <dL/df, v> must equal
  <du, u(f = v)> from the oracle's forward sum, to rounding.  A wrong
  transpose, a dropped filter term or a wrong Re/Im convention fails this;
* the position gradients against central finite differences of the oracle's
  forward sums.  A sign error, a missing E1 derivative or a wrong Gaussian
  derivative fails this;
* the real-space position gradients are translation invariant (every pair
  term depends on x_s - x_t only): sum dL/dpsrc + sum dL/dptar = 0;
* a planted mistake in a restated gradient (the E1 derivative dropped) is
  caught by the same finite-difference check.
"""
import numpy as np
import pytest

from oracle import cref
from tests.conftest import relerr

L = 1.0
XI, NS = 12.0, 3
P, M = 24, 48
W = P * L / M / 2
ETA = (2 * XI * W / (0.95 * np.sqrt(np.pi * P))) ** 2


def _case(seed=3, ns=300, nt=150):
    rng = np.random.default_rng(seed)
    ps = rng.random((ns, 2)) - 0.5
    pt = rng.random((nt, 2)) - 0.5
    f = rng.standard_normal((ns, 2))
    nv = rng.standard_normal((ns, 2))
    nv /= np.linalg.norm(nv, axis=1)[:, None]
    du = rng.standard_normal((nt, 2))
    return ps, pt, f, nv, du


def _fwd(kind, ps, pt, f, nv):
    if kind == "slp_real":
        return cref.slp_real(ps, pt, f, XI, NS, NS, L, L)
    if kind == "dlp_real":
        return cref.dlp_real(ps, pt, f, nv, XI, NS, NS, L, L)
    if kind == "slp_kspace":
        return cref.slp_kspace(ps, pt, XI, ETA, f, M, M, L, L, W, P)
    return cref.dlp_kspace(ps, pt, XI, ETA, f, nv, M, M, L, L, W, P)


def _adj(kind, ps, pt, f, nv, du):
    if kind == "slp_real":
        return cref.slp_real_adj(ps, pt, f, XI, NS, NS, L, L, du)
    if kind == "dlp_real":
        return cref.dlp_real_adj(ps, pt, f, nv, XI, NS, NS, L, L, du)
    if kind == "slp_kspace":
        return cref.slp_kspace_adj(ps, pt, XI, ETA, f, M, M, L, L, W, P, du)
    return cref.dlp_kspace_adj(ps, pt, XI, ETA, f, nv, M, M, L, L, W, P, du)


KINDS = ["slp_real", "dlp_real", "slp_kspace", "dlp_kspace"]


@pytest.mark.parametrize("kind", KINDS)
def test_density_adjoint_identity(kind):
    ps, pt, f, nv, du = _case()
    g = _adj(kind, ps, pt, f, nv, du)
    rng = np.random.default_rng(11)
    v = rng.standard_normal(f.shape)
    lin = np.sum(du * _fwd(kind, ps, pt, v, nv))
    assert abs(np.sum(g["df"] * v) - lin) <= 1e-13 * np.sum(np.abs(g["df"] * v)), kind
    if kind.startswith("dlp"):
        lin = np.sum(du * _fwd(kind, ps, pt, f, v))
        assert abs(np.sum(g["dn"] * v) - lin) <= 1e-13 * np.sum(np.abs(g["dn"] * v)), kind


def _assignment(x):
    """Cell and grid anchor of a point: the piecewise-constant part."""
    cell = tuple(np.floor((x + L / 2) / (L / NS)).astype(int))
    return cell, cref.find_closest_node(x[0], x[1], L, L, L / M, P)[:2]


def _fd_errors(kind, ps, pt, f, nv, du, g, step=1e-6, idx=(3, 17, 42, 99)):
    def loss(a, b):
        return np.sum(du * _fwd(kind, a, b, f, nv))

    for arr in (ps, pt):
        for i in idx:
            for c in range(2):
                for sgn in (1, -1):
                    x = arr[i].copy()
                    x[c] += sgn * step
                    assert _assignment(x) == _assignment(arr[i]), (i, c)
    errs = []
    for side, name in (("s", "dpsrc"), ("t", "dptar")):
        scale = np.abs(g[name]).max()
        for i in idx:
            for c in range(2):
                a = (ps if side == "s" else pt).copy()
                a[i, c] += step
                lp = loss(a, pt) if side == "s" else loss(ps, a)
                a[i, c] -= 2 * step
                lm = loss(a, pt) if side == "s" else loss(ps, a)
                errs.append(abs((lp - lm) / (2 * step) - g[name][i, c]) / scale)
    return max(errs)


@pytest.mark.parametrize("kind", KINDS)
def test_position_gradients_finite_differences(kind):
    # central differences, step 1e-6 (no perturbed point changes cell or
    # anchor): truncation ~ step^2 (third derivative), rounding ~ 1e-16 / step
    ps, pt, f, nv, du = _case()
    g = _adj(kind, ps, pt, f, nv, du)
    assert _fd_errors(kind, ps, pt, f, nv, du, g) < 2e-8, kind


@pytest.mark.parametrize("kind", ["slp_real", "dlp_real"])
def test_real_space_translation_invariance(kind):
    ps, pt, f, nv, du = _case(seed=5)
    g = _adj(kind, ps, pt, f, nv, du)
    tot = g["dpsrc"].sum(axis=0) + g["dptar"].sum(axis=0)
    assert np.abs(tot).max() < 1e-12 * np.abs(g["dpsrc"]).sum(), tot


def test_coincident_points_and_sparse_du():
    # targets = sources: the self pairs carry no position gradient; du = 0 on
    # a target removes it from every gradient
    ps, _, f, nv, du = _case(seed=7, ns=200, nt=200)
    du[::2] = 0.0
    for kind in KINDS:
        g = _adj(kind, ps, ps, f, nv, du)
        sub = np.arange(200)[1::2]
        g2 = _adj(kind, ps, ps[sub], f, nv, du[sub])
        assert relerr(g2["dpsrc"], g["dpsrc"]) < 1e-13, kind
        assert relerr(g2["df"], g["df"]) < 1e-13, kind


def _slp_pair_grad_without_e1_term(d, r2, xi, a, b):
    """Stokeslet pair gradient restated with the E1 derivative dropped (a
    planted mistake): must fail the finite-difference check."""
    e = np.exp(-xi * xi * r2)
    p, q, s = d @ a, d @ b, a @ b
    cd = -2 * xi * xi * e * (p * q / r2 - s) - 2 * p * q * e / r2 ** 2
    return cd * d + e * (q * a + p * b) / r2


def _slp_pair_grad(d, r2, xi, a, b):
    e = np.exp(-xi * xi * r2)
    p, q, s = d @ a, d @ b, a @ b
    cd = -e * s / r2 - 2 * xi * xi * e * (p * q / r2 - s) - 2 * p * q * e / r2 ** 2
    return cd * d + e * (q * a + p * b) / r2


def test_planted_mistake_fails_finite_differences():
    import mpmath

    rng = np.random.default_rng(2)
    xi = 9.0
    a, b = rng.standard_normal(2), rng.standard_normal(2)

    def phi(d):
        r2 = d @ d
        e = np.exp(-xi * xi * r2)
        E = float(mpmath.e1(xi * xi * r2))
        G = 0.5 * E * np.eye(2) + e * (np.outer(d, d) / r2 - np.eye(2))
        return a @ G @ b

    d = np.array([0.031, -0.047])
    h = 1e-7
    fd = np.array([(phi(d + h * np.eye(2)[l]) - phi(d - h * np.eye(2)[l])) / (2 * h) for l in range(2)])
    good = _slp_pair_grad(d, d @ d, xi, a, b)
    bad = _slp_pair_grad_without_e1_term(d, d @ d, xi, a, b)
    assert np.abs(good - fd).max() < 1e-6 * np.abs(fd).max()
    assert np.abs(bad - fd).max() > 1e-3 * np.abs(fd).max()
