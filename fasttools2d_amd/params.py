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
def slp_find_xi(Q, Lx, Ly, rc, tol):
    f = lambda x: math.sqrt(Q * math.pi / (4 * Lx * Ly * x)) * math.exp(-x * x * rc * rc) - tol
    fp = lambda x: (-math.sqrt(Q * math.pi / (4 * Lx * Ly)) * math.exp(-x * x * rc * rc)
                    * (0.5 * x ** -1.5 + 2 * rc * rc * x * math.sqrt(1 / x)))
    return _newton(f, fp, 4 / rc, tol, "xi")


def slp_find_kinfb(Q, Lx, Ly, xi, tol):
    c = 4 * Q * math.pi * max(Lx, Ly) / (Lx ** 3 * Ly ** 3)
    f = lambda k: math.sqrt(c / k) * math.exp(-k * k / (4 * xi * xi)) - tol
    fp = lambda k: (-math.sqrt(c) * math.exp(-k * k / (4 * xi * xi))
                    * (0.5 * k ** -1.5 + math.sqrt(1 / k) * 2 * k / (4 * xi * xi)))
    k = _newton(f, fp, matlab_round(5 * xi), tol, "k")
    return int(matlab_round(Lx * k / (2 * math.pi)))


# ---- stresslet (DLP) estimates, StokesDLP_ewald_2p.m:171-225 ------------------
def dlp_find_xi(Q, Lx, Ly, rc, tol):
    s = math.sqrt(2 * math.pi * Q / (Lx * Ly))
    f = lambda a: math.exp(-a * a * rc * rc) * a * rc * s - tol
    fp = lambda a: (1 - 2 * a * a * rc * rc) * math.exp(-a * a * rc * rc) * s * rc
    return _newton(f, fp, 1 / rc, tol, "xi")


def dlp_find_kinfb(Q, Lx, Ly, xi, tol):
    # MATLAB: sqrt(k*8*pi*Q*max(Lx,Ly)/Lx^3*Ly^3) -- division binds to Lx^3 only
    c = 8 * math.pi * Q * max(Lx, Ly) / Lx ** 3 * Ly ** 3
    f = lambda k: math.sqrt(k * c) * math.exp(-k * k / (4 * xi * xi)) - tol
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
