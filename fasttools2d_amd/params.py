"""Ewald parameter selection and coordinate wrapping of the reference drivers.

Restates the parameter logic of

* PeriodicEwald/src/matlab/StokesSLP_ewald_2p.m:27-104 (defaults, wrap,
  ``rat``, cell count, ``find_xi`` :178-203, ``find_kinfb`` :147-173), and
* PeriodicEwald/src/matlab/StokesDLP_ewald_2p.m:28-125 (same skeleton with the
  stresslet error estimates, ``find_xi`` :202-225, ``find_kinfb`` :171-197).

Both drivers call ``find_kinfb(Q, Lx, Lx, xi, tol)`` (``Lx`` twice), and the
DLP ``find_kinfb`` keeps MATLAB's operator precedence ``.../Lx^3*Ly^3``; both
quirks are reproduced because they decide the grid size.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, asdict

import numpy as np


def matlab_round(x: float) -> float:
    """MATLAB ``round``: halves away from zero."""
    return math.copysign(math.floor(abs(x) + 0.5), x)


def matlab_rat(x: float, tol: float | None = None) -> tuple[int, int]:
    """Continued-fraction rational approximation ``N/D`` of ``x`` as MATLAB's
    ``rat`` computes it (default tolerance ``1e-6*|x|``)."""
    if tol is None:
        tol = 1e-6 * abs(x)
    if not math.isfinite(x):
        raise ValueError("rat needs a finite value")
    y = x
    lastn, lastd = 1.0, 0.0
    n = matlab_round(y)
    d = 1.0
    frac = y - n
    while abs(x - n / d) >= tol:
        flip = 1.0 / frac
        step = matlab_round(flip)
        frac = flip - step
        n, lastn = n * step + lastn, n
        d, lastd = d * step + lastd, d
    if d < 0:
        n, d = -n, -d
    return int(n), int(d)


def wrap(x, L):
    """``mod(x + L/2, L) - L/2`` with MATLAB's ``mod`` (result in [0, L))."""
    x = np.asarray(x, dtype=np.float64)
    return np.mod(x + L / 2, L) - L / 2


def _newton(fun, dfun, x0, tol, what):
    x = x0
    it = 0
    while abs(fun(x)) > tol:
        if it > 100:
            warnings.warn(f"Couldn't find suitable {what}! Break", RuntimeWarning)
            break
        step = min(fun(x) / dfun(x), 0.9 * x)   # limit the step towards 0
        x = x - step
        it += 1
    return x


# ---- Stokeslet (SLP) estimates, StokesSLP_ewald_2p.m:147-203 ------------------
def slp_real_estimate(Q, Lx, Ly, xi, rc):
    """Real-space truncation estimate of StokesSLP_ewald_2p.m:147-176."""
    return math.sqrt(Q * math.pi / (4 * Lx * Ly * xi)) * math.exp(-xi * xi * rc * rc)


def slp_kspace_estimate(Q, Lx, Ly, xi, k):
    """k-space truncation estimate of StokesSLP_ewald_2p.m:178-203 at wavenumber k."""
    c = 4 * Q * math.pi * max(Lx, Ly) / (Lx ** 3 * Ly ** 3)
    return math.sqrt(c / k) * math.exp(-k * k / (4 * xi * xi))


def dlp_real_estimate(Q, Lx, Ly, xi, rc):
    """Real-space truncation estimate of StokesDLP_ewald_2p.m:171-198."""
    return math.exp(-xi * xi * rc * rc) * xi * rc * math.sqrt(2 * math.pi * Q / (Lx * Ly))


def dlp_kspace_estimate(Q, Lx, Ly, xi, k):
    """k-space truncation estimate of StokesDLP_ewald_2p.m:200-225, with MATLAB's
    precedence in sqrt(k*8*pi*Q*max(Lx,Ly)/Lx^3*Ly^3)."""
    c = 8 * math.pi * Q * max(Lx, Ly) / Lx ** 3 * Ly ** 3
    return math.sqrt(k * c) * math.exp(-k * k / (4 * xi * xi))


def slp_find_xi(Q, Lx, Ly, rc, tol):
    f = lambda x: slp_real_estimate(Q, Lx, Ly, x, rc) - tol
    fp = lambda x: (-math.sqrt(Q * math.pi / (4 * Lx * Ly)) * math.exp(-x * x * rc * rc)
                    * (0.5 * x ** -1.5 + 2 * rc * rc * x * math.sqrt(1 / x)))
    return _newton(f, fp, 4 / rc, tol, "xi")


def slp_find_kinfb(Q, Lx, Ly, xi, tol):
    c = 4 * Q * math.pi * max(Lx, Ly) / (Lx ** 3 * Ly ** 3)
    f = lambda k: slp_kspace_estimate(Q, Lx, Ly, xi, k) - tol
    fp = lambda k: (-math.sqrt(c) * math.exp(-k * k / (4 * xi * xi))
                    * (0.5 * k ** -1.5 + math.sqrt(1 / k) * 2 * k / (4 * xi * xi)))
    k = _newton(f, fp, matlab_round(5 * xi), tol, "k")
    return int(matlab_round(Lx * k / (2 * math.pi)))


# ---- stresslet (DLP) estimates, StokesDLP_ewald_2p.m:171-225 ------------------
def dlp_find_xi(Q, Lx, Ly, rc, tol):
    s = math.sqrt(2 * math.pi * Q / (Lx * Ly))
    f = lambda a: dlp_real_estimate(Q, Lx, Ly, a, rc) - tol
    fp = lambda a: (1 - 2 * a * a * rc * rc) * math.exp(-a * a * rc * rc) * s * rc
    return _newton(f, fp, 1 / rc, tol, "xi")


def dlp_find_kinfb(Q, Lx, Ly, xi, tol):
    # MATLAB: sqrt(k*8*pi*Q*max(Lx,Ly)/Lx^3*Ly^3) -- division binds to Lx^3 only
    c = 8 * math.pi * Q * max(Lx, Ly) / Lx ** 3 * Ly ** 3
    f = lambda k: dlp_kspace_estimate(Q, Lx, Ly, xi, k) - tol
    fp = lambda k: (math.sqrt(c) * math.exp(-k * k / (4 * xi * xi))
                    * (0.5 * k ** -0.5 - 2 * k ** 1.5 / (4 * xi * xi)))
    k = _newton(f, fp, matlab_round(5 * xi), tol, "k")
    return int(matlab_round(Lx * k / (2 * math.pi)))


@dataclass
class EwaldParams:
    P: int
    Nb: float
    tol: float
    nside_x: int
    nside_y: int
    rc: float
    xi: float
    kinf: int
    Mx: int
    My: int
    w: float
    eta: float
    A: int
    B: int

    def asdict(self):
        return asdict(self)


def default_Nb(npts: int, P: int) -> float:
    """``Nb = min(P*round(log2(npts)), npts/4)`` (StokesSLP_ewald_2p.m:33)."""
    return min(P * matlab_round(math.log2(npts)), npts / 4)


def select_params(kind: str, nsrc: int, ntar: int, f, Lx: float, Ly: float, *, P: int = 24,
                  Nb: float | None = None, tol: float = 1e-16, M: int | None = None):
    """Parameters the reference driver would pick (``kind`` = 'slp' or 'dlp').

    ``M`` overrides the driver's grid ``min(2*kinf, 10000)`` (both ``Mx`` and
    ``My`` scale with the box aspect ``A:B`` as in the driver); ``w`` and
    ``eta`` follow the grid.
    """
    npts = nsrc + ntar
    if Nb is None:
        Nb = default_Nb(npts, P)
    A, B = matlab_rat(Lx / Ly)
    f = np.asarray(f, dtype=np.float64)
    Q = float(np.sum(f * f)) + 1.0
    a = math.ceil(math.sqrt(npts / (Nb * A * B)))
    m = 0.95 * math.sqrt(math.pi * P)
    nside_x, nside_y = a * A, a * B
    rc = Lx / nside_x
    if kind == "slp":
        xi = slp_find_xi(Q, Lx, Ly, rc, tol)
        kinf = slp_find_kinfb(Q, Lx, Lx, xi, tol)
    elif kind == "dlp":
        xi = dlp_find_xi(Q, Lx, Ly, rc, tol)
        kinf = dlp_find_kinfb(Q, Lx, Lx, xi, tol)
    else:
        raise ValueError("kind must be 'slp' or 'dlp'")
    Mb = min(2 * kinf, 10000) if M is None else int(M)
    My = B * Mb
    Mx = A * Mb
    w = P * Lx / Mx / 2
    eta = (2 * xi * w / m) ** 2
    return EwaldParams(P=P, Nb=Nb, tol=tol, nside_x=nside_x, nside_y=nside_y, rc=rc, xi=xi,
                       kinf=kinf, Mx=Mx, My=My, w=w, eta=eta, A=A, B=B)


# ---- fixed-grid mode -----------------------------------------------------------
# The drivers pick the cell size first (Nb points per cell), then xi from the
# real-space estimate and the grid from the k-space estimate.  With the grid
# given instead, the same two estimates are inverted the other way round: xi
# so that the k-space estimate at the grid's k_inf = M/2 equals tol, then the
# smallest cutoff rc whose real-space estimate at that xi is tol, and the
# largest cell count whose cells are at least rc wide.
def slp_xi_from_kinf(Q, Lx, kinf, tol):
    """Inverse of the Stokeslet k-space estimate (StokesSLP_ewald_2p.m:178-203,
    with Lx passed for both sides as the driver does) at k = 2 pi kinf / Lx."""
    k = 2 * math.pi * kinf / Lx
    c = 4 * Q * math.pi * max(Lx, Lx) / (Lx ** 3 * Lx ** 3)
    return k / (2 * math.sqrt(math.log(math.sqrt(c / k) / tol)))


def dlp_xi_from_kinf(Q, Lx, kinf, tol):
    """Inverse of the stresslet k-space estimate (StokesDLP_ewald_2p.m:200-225,
    the driver's Lx-twice and precedence forms) at k = 2 pi kinf / Lx."""
    k = 2 * math.pi * kinf / Lx
    c = 8 * math.pi * Q * max(Lx, Lx) / Lx ** 3 * Lx ** 3
    return k / (2 * math.sqrt(math.log(math.sqrt(k * c) / tol)))


def slp_rc_from_xi(Q, Lx, Ly, xi, tol):
    """Inverse of the Stokeslet real-space estimate (StokesSLP_ewald_2p.m:147-176)."""
    return math.sqrt(math.log(math.sqrt(Q * math.pi / (4 * Lx * Ly * xi)) / tol)) / xi


def dlp_rc_from_xi(Q, Lx, Ly, xi, tol):
    """Inverse of the stresslet real-space estimate (StokesDLP_ewald_2p.m:171-198):
    x e^{-x^2} s = tol with x = xi rc, s = sqrt(2 pi Q / (Lx Ly)), on the
    decreasing branch x > 1/sqrt(2) (Newton on x^2 - ln x = ln(s / tol))."""
    b = math.log(math.sqrt(2 * math.pi * Q / (Lx * Ly)) / tol)
    x = math.sqrt(max(b, 0.5)) + 1.0
    for _ in range(100):
        g = x * x - math.log(x) - b
        x_new = x - g / (2 * x - 1 / x)
        if abs(x_new - x) <= 1e-15 * x:
            x = x_new
            break
        x = max(x_new, 0.75)
    return x / xi


def select_params_grid(kind: str, nsrc: int, ntar: int, f, Lx: float, Ly: float, *, M: int,
                       P: int = 24, tol: float = 1e-16):
    """Parameters for a given grid (fixed-grid mode; no reference counterpart,
    built from the drivers' own estimates).  ``M`` is the base grid (Mx = A*M,
    My = B*M for the box aspect A:B, as in the drivers); xi comes from the
    k-space estimate at k_inf = M/2, the cutoff from the real-space estimate
    at that xi, and nside_x = a*A, nside_y = a*B with the largest a whose cells
    are at least the cutoff wide.  ``Nb`` reports the resulting points per
    cell."""
    M = int(M)
    if M < 2 or M % 2:
        raise ValueError("M must be even and >= 2")
    A, B = matlab_rat(Lx / Ly)
    f = np.asarray(f, dtype=np.float64)
    Q = float(np.sum(f * f)) + 1.0
    kinf = M // 2
    if kind == "slp":
        xi = slp_xi_from_kinf(Q, Lx, kinf, tol)
        rc_min = slp_rc_from_xi(Q, Lx, Ly, xi, tol)
    elif kind == "dlp":
        xi = dlp_xi_from_kinf(Q, Lx, kinf, tol)
        rc_min = dlp_rc_from_xi(Q, Lx, Ly, xi, tol)
    else:
        raise ValueError("kind must be 'slp' or 'dlp'")
    a = max(1, int(math.floor(Lx / (A * rc_min) * (1 + 1e-12))))
    nside_x, nside_y = a * A, a * B
    rc = Lx / nside_x
    m = 0.95 * math.sqrt(math.pi * P)
    Mx, My = A * M, B * M
    w = P * Lx / Mx / 2
    eta = (2 * xi * w / m) ** 2
    Nb = (nsrc + ntar) / (nside_x * nside_y)
    return EwaldParams(P=P, Nb=Nb, tol=tol, nside_x=nside_x, nside_y=nside_y, rc=rc, xi=xi,
                       kinf=kinf, Mx=Mx, My=My, w=w, eta=eta, A=A, B=B)
