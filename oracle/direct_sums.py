"""Closed-form direct Ewald sums (test infrastructure only).

Vectorised numpy restatements of the reference's MATLAB direct sums, used to
pin the C oracle and the HIP path at truncation tolerance:

* ``slp_real_ds``   -- PeriodicEwald/tests/matlab_sums/stokes_slp_real_ds.m:26-69
* ``dlp_real_ds``   -- PeriodicEwald/tests/matlab_sums/stokes_dlp_real_ds.m:28-76
* ``slp_kspace_ds`` -- PeriodicEwald/tests/matlab_sums/stokes_slp_kspace_ds.m:27-74
* ``dlp_kspace_ds`` -- PeriodicEwald/tests/matlab_sums/stokes_dlp_kspace_ds.m:29-85

The real-space sums visit the 3x3 periodic images with no cutoff; the Fourier
sums visit wavenumbers ``j in (-kinf, kinf]^2 \\ {0}`` without gridding or
FFTs.  ``E1`` is ``scipy.special.exp1``.  Points are ``(N, 2)`` arrays.
"""
from __future__ import annotations

import numpy as np
from scipy.special import exp1

EULER_GAMMA = 0.57721566490153286061


def _split(p):
    p = np.asarray(p, dtype=np.float64)
    return p[:, 0], p[:, 1]


def slp_real_ds(psrc, ptar, f, Lx, Ly, xi):
    xs, ys = _split(psrc)
    xt, yt = _split(ptar)
    f1, f2 = _split(f)
    u1 = np.zeros(len(xt))
    u2 = np.zeros(len(xt))
    for jx in (-1, 0, 1):
        for jy in (-1, 0, 1):
            r1 = (xs[None, :] + jx * Lx) - xt[:, None]
            r2 = (ys[None, :] + jy * Ly) - yt[:, None]
            rr = r1 * r1 + r2 * r2
            r = np.sqrt(rr)
            near = np.abs(r) < 1e-13
            rs = np.where(near, 1.0, rr)
            e = np.exp(-xi * xi * rs)
            E = exp1(xi * xi * rs)
            rdotf = r1 * f1[None, :] + r2 * f2[None, :]
            t1 = e * (r1 * rdotf / rs - f1[None, :]) + f1[None, :] * E / 2
            t2 = e * (r2 * rdotf / rs - f2[None, :]) + f2[None, :] * E / 2
            selfc = -(EULER_GAMMA / 2 + np.log(xi) + 1)
            t1 = np.where(near, selfc * f1[None, :], t1)
            t2 = np.where(near, selfc * f2[None, :], t2)
            u1 += t1.sum(axis=1)
            u2 += t2.sum(axis=1)
    return np.stack([u1, u2], axis=1) / (4 * np.pi)


def dlp_real_ds(psrc, ptar, f, n, Lx, Ly, xi):
    xs, ys = _split(psrc)
    xt, yt = _split(ptar)
    f1, f2 = _split(f)
    n1, n2 = _split(n)
    u1 = np.zeros(len(xt))
    u2 = np.zeros(len(xt))
    for jx in (-1, 0, 1):
        for jy in (-1, 0, 1):
            r1 = (xs[None, :] + jx * Lx) - xt[:, None]
            r2 = (ys[None, :] + jy * Ly) - yt[:, None]
            rr = r1 * r1 + r2 * r2
            near = np.sqrt(rr) < 1e-13
            rs = np.where(near, 1.0, rr)
            rdotf = r1 * f1[None, :] + r2 * f2[None, :]
            rdotn = r1 * n1[None, :] + r2 * n2[None, :]
            ndotf = (n1 * f1 + n2 * f2)[None, :]
            e = np.exp(-xi * xi * rs)
            a = -4 * rdotn * rdotf / (rs * rs) * (1 + xi * xi * rs)
            t1 = e * (a * r1 + 2 * xi * xi * (f1[None, :] * rdotn + n1[None, :] * rdotf + ndotf * r1))
            t2 = e * (a * r2 + 2 * xi * xi * (f2[None, :] * rdotn + n2[None, :] * rdotf + ndotf * r2))
            u1 += np.where(near, 0.0, t1).sum(axis=1)
            u2 += np.where(near, 0.0, t2).sum(axis=1)
    return np.stack([u1, u2], axis=1) / (4 * np.pi)


def _kgrid(Lx, Ly, kinf):
    j = np.arange(-kinf + 1, kinf + 1)
    k1 = 2 * np.pi / Lx * j
    k2 = 2 * np.pi / Ly * j
    K1, K2 = np.meshgrid(k1, k2, indexing="ij")
    K1 = K1.ravel()
    K2 = K2.ravel()
    keep = ~((K1 == 0) & (K2 == 0))
    return K1[keep], K2[keep]


def slp_kspace_ds(psrc, ptar, f, Lx, Ly, xi, kinf):
    xs, ys = _split(psrc)
    xt, yt = _split(ptar)
    f1, f2 = _split(f)
    k1, k2 = _kgrid(Lx, Ly, kinf)
    K2 = k1 * k1 + k2 * k2
    g = (1 + K2 / (4 * xi * xi)) * np.exp(-K2 / (4 * xi * xi)) / K2
    # structure factors S_c(k) = sum_s f_c e^{+i k.x_s}
    ph = np.exp(1j * (np.outer(k1, xs) + np.outer(k2, ys)))
    S1 = ph @ f1
    S2 = ph @ f2
    kdotS = k1 * S1 + k2 * S2
    A1 = (S1 - k1 * kdotS / K2) * g
    A2 = (S2 - k2 * kdotS / K2) * g
    pt = np.exp(-1j * (np.outer(xt, k1) + np.outer(yt, k2)))
    u1 = np.real(pt @ A1) / (Lx * Ly)
    u2 = np.real(pt @ A2) / (Lx * Ly)
    return np.stack([u1, u2], axis=1)


def dlp_kspace_ds(psrc, ptar, f, n, Lx, Ly, xi, kinf, zero_mode=True):
    """Includes the k = 0 mode (stokes_dlp_kspace_ds.m:67-69) unless
    ``zero_mode`` is False; the MEX kernel itself leaves it out."""
    xs, ys = _split(psrc)
    xt, yt = _split(ptar)
    f1, f2 = _split(f)
    n1, n2 = _split(n)
    k1, k2 = _kgrid(Lx, Ly, kinf)
    K2 = k1 * k1 + k2 * k2
    g = (1 + K2 / (4 * xi * xi)) * np.exp(-K2 / (4 * xi * xi)) / K2
    ph = np.exp(1j * (np.outer(k1, xs) + np.outer(k2, ys)))
    # the kernel is bilinear in (f, n): expand into the four products
    S11 = ph @ (f1 * n1)
    S12 = ph @ (f1 * n2)
    S21 = ph @ (f2 * n1)
    S22 = ph @ (f2 * n2)
    # kdotf*kdotn summed over sources: sum_jl k_j k_l S_jl
    kk = k1 * k1 * S11 + k1 * k2 * (S12 + S21) + k2 * k2 * S22
    # f1*kdotn + n1*kdotf + k1*fdotn
    A1 = (k1 * S11 + k2 * S12) + (k1 * S11 + k2 * S21) + k1 * (S11 + S22) - 2 * k1 * kk / K2
    A2 = (k1 * S21 + k2 * S22) + (k1 * S12 + k2 * S22) + k2 * (S11 + S22) - 2 * k2 * kk / K2
    pt = np.exp(-1j * (np.outer(xt, k1) + np.outer(yt, k2)))
    u1 = np.real(pt @ (1j * A1 * g)) / (Lx * Ly)
    u2 = np.real(pt @ (1j * A2 * g)) / (Lx * Ly)
    u = np.stack([u1, u2], axis=1)
    if zero_mode:
        fn = f1 * n1 + f2 * n2
        u[:, 0] -= np.sum(fn * xs) / (Lx * Ly)
        u[:, 1] -= np.sum(fn * ys) / (Lx * Ly)
    return u


# -- the same sums for large source sets ---------------------------------------
# The Fourier sums above form phase matrices of size (#wavenumbers x N).  The
# versions below use that e^{i k.x} = e^{i k1 x} e^{i k2 y} on the wavenumber
# lattice: the structure factors S(j1, j2) = sum_s q_s e^{i k1 x_s} e^{i k2 y_s}
# are the matrix product E1^T diag(q) E2 of per-axis phase matrices, evaluated
# in chunks of sources; the target sums factor the same way.  Same sums, same
# wavenumbers (j in (-kinf, kinf]^2 without 0); only the evaluation order
# differs.

def _axis_phases(x, L, kinf, sign):
    j = np.arange(-kinf + 1, kinf + 1)
    return np.exp(sign * 1j * np.outer(x, 2 * np.pi / L * j))


def _structure_factors(xs, ys, q, Lx, Ly, kinf, chunk=8192):
    """S[c][j1, j2] = sum_s q[c][s] e^{i (k1 x_s + k2 y_s)} for each row c of q."""
    nq = q.shape[0]
    m = 2 * kinf
    S = np.zeros((nq, m, m), dtype=np.complex128)
    for s0 in range(0, len(xs), chunk):
        s1 = min(len(xs), s0 + chunk)
        E1 = _axis_phases(xs[s0:s1], Lx, kinf, +1)        # (c, m)
        E2 = _axis_phases(ys[s0:s1], Ly, kinf, +1)        # (c, m)
        B = np.concatenate([q[c, s0:s1, None] * E2 for c in range(nq)], axis=1)
        S += (E1.T @ B).reshape(m, nq, m).transpose(1, 0, 2)
    return S


def _target_sum(xt, yt, C, Lx, Ly, kinf):
    """Re sum_{j1, j2} C[j1, j2] e^{-i (k1 x_t + k2 y_t)} for every target."""
    T1 = _axis_phases(xt, Lx, kinf, -1)
    T2 = _axis_phases(yt, Ly, kinf, -1)
    return np.real(np.sum((T1 @ C) * T2, axis=1))


def _lattice(Lx, Ly, kinf):
    j = np.arange(-kinf + 1, kinf + 1)
    K1, K2 = np.meshgrid(2 * np.pi / Lx * j, 2 * np.pi / Ly * j, indexing="ij")
    K2s = K1 * K1 + K2 * K2
    zero = K2s == 0
    K2s[zero] = 1.0   # the k = 0 term is dropped below
    return K1, K2, K2s, zero


def slp_kspace_ds_sep(psrc, ptar, f, Lx, Ly, xi, kinf):
    """slp_kspace_ds by per-axis phase matrices (large source sets)."""
    xs, ys = _split(psrc)
    xt, yt = _split(ptar)
    f1, f2 = _split(f)
    k1, k2, K2, zero = _lattice(Lx, Ly, kinf)
    g = (1 + K2 / (4 * xi * xi)) * np.exp(-K2 / (4 * xi * xi)) / K2
    g[zero] = 0.0
    S1, S2 = _structure_factors(xs, ys, np.stack([f1, f2]), Lx, Ly, kinf)
    kdotS = k1 * S1 + k2 * S2
    A1 = (S1 - k1 * kdotS / K2) * g
    A2 = (S2 - k2 * kdotS / K2) * g
    u1 = _target_sum(xt, yt, A1, Lx, Ly, kinf) / (Lx * Ly)
    u2 = _target_sum(xt, yt, A2, Lx, Ly, kinf) / (Lx * Ly)
    return np.stack([u1, u2], axis=1)


def dlp_kspace_ds_sep(psrc, ptar, f, n, Lx, Ly, xi, kinf, zero_mode=True):
    """dlp_kspace_ds by per-axis phase matrices (large source sets)."""
    xs, ys = _split(psrc)
    xt, yt = _split(ptar)
    f1, f2 = _split(f)
    n1, n2 = _split(n)
    k1, k2, K2, zero = _lattice(Lx, Ly, kinf)
    g = (1 + K2 / (4 * xi * xi)) * np.exp(-K2 / (4 * xi * xi)) / K2
    g[zero] = 0.0
    S11, S12, S21, S22 = _structure_factors(
        xs, ys, np.stack([f1 * n1, f1 * n2, f2 * n1, f2 * n2]), Lx, Ly, kinf)
    kk = k1 * k1 * S11 + k1 * k2 * (S12 + S21) + k2 * k2 * S22
    A1 = (k1 * S11 + k2 * S12) + (k1 * S11 + k2 * S21) + k1 * (S11 + S22) - 2 * k1 * kk / K2
    A2 = (k1 * S21 + k2 * S22) + (k1 * S12 + k2 * S22) + k2 * (S11 + S22) - 2 * k2 * kk / K2
    u1 = _target_sum(xt, yt, 1j * A1 * g, Lx, Ly, kinf) / (Lx * Ly)
    u2 = _target_sum(xt, yt, 1j * A2 * g, Lx, Ly, kinf) / (Lx * Ly)
    u = np.stack([u1, u2], axis=1)
    if zero_mode:
        fn = f1 * n1 + f2 * n2
        u[:, 0] -= np.sum(fn * xs) / (Lx * Ly)
        u[:, 1] -= np.sum(fn * ys) / (Lx * Ly)
    return u


def real_ds_chunked(fn, psrc, ptar, *args, chunk=16):
    """fn (slp_real_ds or dlp_real_ds) over targets in chunks (bounded memory)."""
    out = [fn(psrc, ptar[t0:t0 + chunk], *args) for t0 in range(0, len(ptar), chunk)]
    return np.concatenate(out, axis=0)
