"""The reference's invariants of the whole method, run through the HIP path
(the driver entry points, se2p_slp_ewald / se2p_dlp_ewald), on the inputs of
PeriodicEwald/tests/consistency_test.m: 10000 sources and 10000 separate
targets uniform in a 1 x 2 box, densities 10 U[0,1), normals (n1, sqrt(1 - n1^2)).

* The result does not depend on the points per cell Nb (consistency_test.m:
  44-66 and 116-138): Nb = 3, 9, 27 change the cells, xi and the grid.
* Nor, in the fixed-grid mode, on the grid: three grids, each with its own xi
  and cells at the same tolerance.
* Replicating the box (sources copied one period over, Lx doubled) does not
  change the result (:68-84); for the stresslet after subtracting the zero
  mode sum((f.n) x) / (Lx Ly), which depends on the source positions
  (:146-179).

Bar: 1e-10 in the max norm relative to the largest output, at tol 1e-13."""
import numpy as np
import pytest

from tests.conftest import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _inputs(seed):
    rng = np.random.default_rng(seed)
    Lx, Ly = 1.0, 2.0
    f = 10 * rng.random((10000, 2))
    n1 = rng.random(10000)
    nv = np.stack([n1, np.sqrt(1 - n1 ** 2)], 1)
    xs = rng.random((10000, 2)) * [Lx, Ly]
    xt = rng.random((10000, 2)) * [Lx, Ly]
    return xs, xt, f, nv, Lx, Ly


def _run(pe, kind, xs, xt, f, nv, Lx, Ly, **kw):
    if kind == "slp":
        u1, u2, *_ = pe.StokesSLP_ewald_2p(xs[:, 0], xs[:, 1], xt[:, 0], xt[:, 1], f[:, 0],
                                           f[:, 1], Lx, Ly, tol=TOL, **kw)
    else:
        u1, u2, *_ = pe.StokesDLP_ewald_2p(xs[:, 0], xs[:, 1], xt[:, 0], xt[:, 1], nv[:, 0],
                                           nv[:, 1], f[:, 0], f[:, 1], Lx, Ly, tol=TOL, **kw)
    return np.stack([u1.cpu().numpy(), u2.cpu().numpy()], 1)


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_independent_of_points_per_cell(gpu, kind):
    xs, xt, f, nv, Lx, Ly = _inputs(61)
    us = []
    for Nb in (3, 9, 27):
        print(f"[invariants {kind}] Nb={Nb}", flush=True)
        us.append(_run(gpu, kind, xs, xt, f, nv, Lx, Ly, Nb=Nb))
    errs = [relerr(u, us[0]) for u in us[1:]]
    print(f"[invariants {kind}] Nb independence: {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-10


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_independent_of_the_grid_in_fixed_grid_mode(gpu, kind):
    xs, xt, f, nv, Lx, Ly = _inputs(62)
    us = []
    for M in (96, 128, 192):
        print(f"[invariants {kind}] fixed grid M={M}", flush=True)
        us.append(_run(gpu, kind, xs, xt, f, nv, Lx, Ly, M=M, fixed_grid=True))
    errs = [relerr(u, us[0]) for u in us[1:]]
    print(f"[invariants {kind}] grid independence: {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-10


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_periodic_replication(gpu, kind):
    xs, xt, f, nv, Lx, Ly = _inputs(63)
    print(f"[invariants {kind}] replication", flush=True)
    u1 = _run(gpu, kind, xs, xt, f, nv, Lx, Ly)
    xs2 = np.concatenate([xs, xs + [Lx, 0.0]])
    f2, n2 = np.concatenate([f, f]), np.concatenate([nv, nv])
    u2 = _run(gpu, kind, xs2, xt, f2, n2, 2 * Lx, Ly)
    if kind == "dlp":
        # the stresslet's zero mode depends on where the sources are: take it off
        # both results as consistency_test.m:146-179 does (sign: u = -(ur + uk))
        def zero_mode(x, ff, nn, LxLy):
            fn = ff[:, 0] * nn[:, 0] + ff[:, 1] * nn[:, 1]
            return np.array([np.sum(fn * x[:, 0]), np.sum(fn * x[:, 1])]) / LxLy
        u1 = u1 - zero_mode(xs, f, nv, Lx * Ly)
        u2 = u2 - zero_mode(xs2, f2, n2, 2 * Lx * Ly)
    err = relerr(u2, u1)
    print(f"[invariants {kind}] replication: {err:.2e}", flush=True)
    assert err <= 1e-10
