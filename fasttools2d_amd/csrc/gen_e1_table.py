#!/usr/bin/env python3
"""Generate ``e1_table.h``: piecewise polynomial coefficients for the
exponential integral E1(x) used by the real-space Stokeslet kernel.

Method (Chebyshev interpolation, Trefethen "Approximation Theory and
Approximation Practice" ch. 2-4): on each piece [a, b] the target function is
interpolated at NCOEF Chebyshev points of the second kind in 60-digit
arithmetic (mpmath), converted to the monomial basis in the local variable
s = (2x - a - b)/(b - a) in [-1, 1], and rounded to double.

Pieces and target functions:
  * x in [0, 1): two pieces [0, 1/2), [1/2, 1) of
        S(x) = (E1(x) + gamma + ln x) / x = sum_{k>=1} (-1)^{k+1} x^{k-1} / (k k!)
    (an entire function), so that E1(x) = x S(x) - gamma - ln x.
  * x in [1, 1024): 40 pieces, four per octave, split linearly in the mantissa
    (x = 2^e m, m in [1,2), piece 2 + 4e + floor(4(m-1))), of
        g(x) = e^x E1(x),   so that E1(x) = e^{-x} g(x).
  * x >= 1024: E1(x) < 1e-447 underflows; a final all-zero piece.

Run ``python gen_e1_table.py`` to rewrite ``e1_table.h``; the script also
prints the worst relative error of the double-precision evaluation.
"""
from __future__ import annotations

import os

import mpmath as mp

mp.mp.dps = 60
NCOEF = 15            # degree 14
NPIECE = 2 + 40 + 1   # small-x pieces, quarter-octave pieces, zero piece


def S_small(x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1)
    return (mp.e1(x) + mp.euler + mp.log(x)) / x


def g_large(x):
    x = mp.mpf(x)
    return mp.exp(x) * mp.e1(x)


def pieces():
    out = [(mp.mpf(0), mp.mpf("0.5"), S_small), (mp.mpf("0.5"), mp.mpf(1), S_small)]
    for e in range(10):
        for q in range(4):
            a = mp.mpf(2) ** e * (1 + mp.mpf(q) / 4)
            b = mp.mpf(2) ** e * (1 + mp.mpf(q + 1) / 4)
            out.append((a, b, g_large))
    return out


def cheb_to_monomial(a, b, fn, n):
    # values at Chebyshev points of the second kind
    nodes = [mp.cos(mp.pi * j / (n - 1)) for j in range(n)]
    vals = [fn((a + b) / 2 + (b - a) / 2 * s) for s in nodes]
    # Chebyshev coefficients by the discrete cosine sum
    c = []
    for k in range(n):
        acc = mp.mpf(0)
        for j in range(n):
            wj = mp.mpf("0.5") if j in (0, n - 1) else mp.mpf(1)
            acc += wj * vals[j] * mp.cos(mp.pi * k * j / (n - 1))
        acc *= mp.mpf(2) / (n - 1)
        if k in (0, n - 1):
            acc /= 2
        c.append(acc)
    # T_k -> monomials
    T = [[mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]]
    for k in range(2, n):
        prev, prev2 = T[k - 1], T[k - 2]
        cur = [mp.mpf(0)] * (k + 1)
        for i, v in enumerate(prev):
            cur[i + 1] += 2 * v
        for i, v in enumerate(prev2):
            cur[i] -= v
        T.append(cur)
    mono = [mp.mpf(0)] * n
    for k in range(n):
        for i, v in enumerate(T[k]):
            mono[i] += c[k] * v
    return mono


def horner(coef, s):
    acc = 0.0
    for cf in reversed(coef):
        acc = acc * s + cf
    return acc


def main():
    rows = []
    worst = 0.0
    for a, b, fn in pieces():
        mono = cheb_to_monomial(a, b, fn, NCOEF)
        coef = [float(v) for v in mono]
        sa = float(2 / (b - a))
        sb = float(-(a + b) / (b - a))
        # check the double evaluation of the full E1 against mpmath
        for t in range(0, 401):
            x = a + (b - a) * mp.mpf(t) / 400
            if x == 0:
                continue
            xd = float(x)
            s = xd * sa + sb
            v = horner(coef, s)
            if fn is S_small:
                import math
                e1d = xd * v - 0.57721566490153286061 - math.log(xd)
            else:
                import math
                e1d = math.exp(-xd) * v
            ref = mp.e1(mp.mpf(xd))
            if ref != 0 and xd < 690:
                worst = max(worst, abs(float((mp.mpf(e1d) - ref) / ref)))
        rows.append((sa, sb, coef))
    rows.append((0.0, 0.0, [0.0] * NCOEF))
    assert len(rows) == NPIECE
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "e1_table.h"), "w") as fh:
        fh.write("// Generated by gen_e1_table.py -- do not edit.\n")
        fh.write("// Piecewise Chebyshev-interpolant coefficients for E1(x); see the generator.\n")
        fh.write(f"// Worst relative error of the double evaluation on the check grid: {worst:.3e}\n")
        fh.write("#pragma once\n")
        fh.write(f"#define SE2P_E1_NCOEF {NCOEF}\n#define SE2P_E1_NPIECE {NPIECE}\n")
        fh.write("// per piece: s = x*scale + shift, then sum_k c_k s^k\n")
        fh.write("static const double se2p_e1_scale_shift[SE2P_E1_NPIECE][2] = {\n")
        for sa, sb, _ in rows:
            fh.write(f"  {{{sa!r}, {sb!r}}},\n")
        fh.write("};\n")
        fh.write("static const double se2p_e1_coef[SE2P_E1_NPIECE][SE2P_E1_NCOEF] = {\n")
        for _, _, coef in rows:
            fh.write("  {" + ", ".join(repr(c) for c in coef) + "},\n")
        fh.write("};\n")
    print(f"wrote e1_table.h; worst relative error {worst:.3e}")


if __name__ == "__main__":
    main()
