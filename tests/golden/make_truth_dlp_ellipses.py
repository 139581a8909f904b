"""Write truth_dlp_ellipses_N1e5.json: the FP64 direct Ewald sum of the
stresslet (real-space plus Fourier-space parts, the two MEX outputs added) on
a 256-target subset of the benchmark workload.

Workload: 200 non-overlapping ellipses x 500 nodes (N = 1e5, seed 2), unit
box, Ewald splitting xi from the driver formulas at tol 1e-10 (grid 512^2).
Targets: numpy default_rng(55).choice(N, 256, replace=False), the same subset
the benchmark checks.

Truth = dlp_real_ds (all 3x3 periodic images, no cutoff) + dlp_kspace_ds_sep
(every wavenumber j in (-256, 256]^2 except 0, without the DLP zero mode,
which the MEX kernel leaves to the driver); oracle/direct_sums.py restates
PeriodicEwald/tests/matlab_sums/stokes_dlp_{real,kspace}_ds.m.  At
xi = 110, e^{-K^2/(4 xi^2)} at |j| = 256 is below 1e-23 and e^{-xi^2 r^2} at
half a box is below 1e-1300, so both truncations are far below FP64 rounding.

Runs on the CPU in about a minute:  python tests/golden/make_truth_dlp_ellipses.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fasttools2d_amd import workloads as W  # noqa: E402
from oracle import direct_sums as ds  # noqa: E402

KINF = 256


def main():
    w, p = W.dlp_ellipses_config(seed=2, n_ell=200, n_nodes=500, tol=1e-10, M=512)
    N = len(w.psrc)
    sub = np.random.default_rng(55).choice(N, 256, replace=False)
    pt = w.ptar[sub]
    ur = ds.real_ds_chunked(ds.dlp_real_ds, w.psrc, pt, w.f, w.n, w.Lx, w.Ly, p.xi)
    uk = ds.dlp_kspace_ds_sep(w.psrc, pt, w.f, w.n, w.Lx, w.Ly, p.xi, KINF, zero_mode=False)
    out = {
        "workload": "stresslet DLP, 200 ellipses x 500 nodes (N=1e5), seed 2, unit box",
        "xi": p.xi,
        "kinf": KINF,
        "targets": [int(i) for i in sub],
        "u_real": ur.tolist(),
        "u_kspace": uk.tolist(),
        "generator": "tests/golden/make_truth_dlp_ellipses.py",
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "truth_dlp_ellipses_N1e5.json")
    with open(path, "w") as fh:
        json.dump(out, fh)
    print(f"wrote {path}: max |u| = {np.abs(ur + uk).max():.6e}")


if __name__ == "__main__":
    main()
