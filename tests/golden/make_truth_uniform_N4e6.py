"""Write truth_uniform_N4e6.json: a tighter Ewald sum of the largest workload
on 256 targets, to measure the error of the runs on its stated 2048^2 grid.

Workload: N = 4e6 uniform points in the unit box (fasttools2d_amd.workloads.
uniform_box(4_000_000, seed=5)), targets = sources, N(0,1) densities, unit
normals at uniform angles.  Targets: numpy default_rng(5).choice(N, 256,
replace=False), the subset the at-scale GPU test and the benchmark check.

Truth = the oracle's real-space sum plus its Fourier-space sum (the two MEX
outputs added; the stresslet without the zero mode, which the MEX kernel
leaves to the driver) at tol 1e-14: the drivers' default cell size (nside
121), their xi at that tolerance, and a 4096^2 grid (at least the drivers'
2 k_inf of 2810 and 3396; a power of two for the oracle's FFT).  The measured
runs use tol 1e-12 on 2048^2 with xi and the cells from the fixed-grid mode
(nside 99 and 80), so no parameter is shared.  The total is independent of
xi up to the truncation errors.

Runs on the CPU (OpenMP) in a few minutes:
    python tests/golden/make_truth_uniform_N4e6.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fasttools2d_amd import params as Pm  # noqa: E402
from fasttools2d_amd import workloads as W  # noqa: E402
from oracle import cref  # noqa: E402

N = 4_000_000
TOL = 1e-14
M = 4096


def main():
    w = W.uniform_box(N, seed=5)
    sub = np.random.default_rng(5).choice(N, 256, replace=False)
    pt = w.ptar[sub]
    cref.set_threads(os.cpu_count() or 1)
    out = {"workload": "N=4e6 uniform points, unit box, seed 5, targets = sources",
           "targets": [int(i) for i in sub], "tol": TOL, "generator":
           "tests/golden/make_truth_uniform_N4e6.py"}
    for kind in ("slp", "dlp"):
        p = Pm.select_params(kind, N, N, w.f, 1.0, 1.0, tol=TOL, M=M)
        t0 = time.time()
        if kind == "slp":
            ur = cref.slp_real(w.psrc, pt, w.f, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
            uk = cref.slp_kspace(w.psrc, pt, p.xi, p.eta, w.f, p.Mx, p.My, 1.0, 1.0, p.w, p.P)
        else:
            ur = cref.dlp_real(w.psrc, pt, w.f, w.n, p.xi, p.nside_x, p.nside_y, 1.0, 1.0)
            uk = cref.dlp_kspace(w.psrc, pt, p.xi, p.eta, w.f, w.n, p.Mx, p.My, 1.0, 1.0, p.w, p.P)
        out[kind] = {"xi": p.xi, "nside": p.nside_x, "M": p.Mx, "eta": p.eta,
                     "u": (ur + uk).tolist()}
        print(f"{kind}: nside {p.nside_x} xi {p.xi:.4f} M {p.Mx} in {time.time() - t0:.1f} s, "
              f"max |u| = {np.abs(ur + uk).max():.6e}", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "truth_uniform_N4e6.json")
    with open(path, "w") as fh:
        json.dump(out, fh)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
