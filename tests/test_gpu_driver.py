"""The driver entry points (se2p_slp_ewald / se2p_dlp_ewald): parameters,
wrapping and composition of StokesSLP_ewald_2p.m / StokesDLP_ewald_2p.m on the
GPU, against the oracle composed the same way, plus the stresslet identity on
the product."""
import numpy as np
import pytest
import torch

from oracle import cref
from fasttools2d_amd import params as Pm
from fasttools2d_amd import workloads as W
from tests.conftest import relerr

pytestmark = pytest.mark.gpu


def _oracle_full(kind, xs_u, xt_u, f, n, Lx, Ly, prm):
    ps = np.stack([Pm.wrap(xs_u[:, 0], Lx), Pm.wrap(xs_u[:, 1], Ly)], 1)
    pt = np.stack([Pm.wrap(xt_u[:, 0], Lx), Pm.wrap(xt_u[:, 1], Ly)], 1)
    if kind == "slp":
        ur = cref.slp_real(ps, pt, f, prm.xi, prm.nside_x, prm.nside_y, Lx, Ly)
        uk = cref.slp_kspace(ps, pt, prm.xi, prm.eta, f, prm.Mx, prm.My, Lx, Ly, prm.w, prm.P)
        return ur + uk
    ur = cref.dlp_real(ps, pt, f, n, prm.xi, prm.nside_x, prm.nside_y, Lx, Ly)
    uk = cref.dlp_kspace(ps, pt, prm.xi, prm.eta, f, n, prm.Mx, prm.My, Lx, Ly, prm.w, prm.P)
    fn = f[:, 0] * n[:, 0] + f[:, 1] * n[:, 1]
    uk[:, 0] -= np.sum(fn * xs_u[:, 0]) / (Lx * Ly)
    uk[:, 1] -= np.sum(fn * xs_u[:, 1]) / (Lx * Ly)
    return -(ur + uk)


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_driver_matches_oracle_composition(gpu, kind):
    pe = gpu
    rng = np.random.default_rng(31)
    Lx, Ly = 1.0, 2.0
    xs = rng.uniform(-1.3, 2.2, (800, 2)) * [Lx, Ly]      # unwrapped, several periods away
    xt = rng.uniform(0.0, 1.0, (300, 2)) * [Lx, Ly]
    f = rng.standard_normal((800, 2))
    n = rng.standard_normal((800, 2))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    print(f"[driver {kind}]", flush=True)
    if kind == "slp":
        u1, u2, ur, uk, xi, prm = pe.StokesSLP_ewald_2p(xs[:, 0], xs[:, 1], xt[:, 0], xt[:, 1],
                                                        f[:, 0], f[:, 1], Lx, Ly, tol=1e-10,
                                                        return_params=True)
    else:
        u1, u2, ur, uk, xi, prm = pe.StokesDLP_ewald_2p(xs[:, 0], xs[:, 1], xt[:, 0], xt[:, 1],
                                                        n[:, 0], n[:, 1], f[:, 0], f[:, 1], Lx, Ly,
                                                        tol=1e-10, return_params=True)
    u = np.stack([u1.cpu().numpy(), u2.cpu().numpy()], 1)
    c = _oracle_full(kind, xs, xt, f, n, Lx, Ly, prm)
    assert relerr(u, c) <= 1e-12
    ref = Pm.select_params(kind, 800, 300, f, Lx, Ly, tol=1e-10)
    assert (prm.nside_x, prm.Mx, prm.My) == (ref.nside_x, ref.Mx, ref.My)


@pytest.mark.parametrize("comp", [0, 1])
def test_stresslet_identity_on_the_product(gpu, comp):
    """u = e_comp inside the circle and 0 outside
    (PeriodicEwald/tests/stresslet_identity_test.m:60-78)."""
    pe = gpu
    w = W.stresslet_circle(comp=comp)
    xs = w.psrc_unwrapped
    pt = w.ptar
    print(f"[identity comp={comp}]", flush=True)
    u1, u2, ur, uk, xi = pe.StokesDLP_ewald_2p(xs[:, 0], xs[:, 1], pt[:, 0], pt[:, 1],
                                               w.n[:, 0], w.n[:, 1], w.f[:, 0], w.f[:, 1],
                                               w.Lx, w.Ly, tol=1e-14)
    u = np.stack([u1.cpu().numpy(), u2.cpu().numpy()], 1)
    e = np.zeros(2)
    e[comp] = 1.0
    assert np.abs(u[:16] - e).max() < 1e-10
    assert np.abs(u[16:]).max() < 1e-10


def test_driver_targets_equal_sources(gpu):
    """Targets passed as the same arrays as the sources share the cell lists."""
    pe = gpu
    w = W.uniform_box(1000, seed=1)
    x, y = w.psrc_unwrapped[:, 0], w.psrc_unwrapped[:, 1]
    u1, u2, ur, uk, xi, prm = pe.StokesSLP_ewald_2p(x, y, x, y, w.f[:, 0], w.f[:, 1], 1.0, 1.0,
                                                    tol=1e-10, M=64, return_params=True)
    u = np.stack([u1.cpu().numpy(), u2.cpu().numpy()], 1)
    c = _oracle_full("slp", w.psrc_unwrapped, w.psrc_unwrapped, w.f, None, 1.0, 1.0, prm)
    assert relerr(u, c) <= 1e-12 and prm.Mx == 64
