#!/usr/bin/env python3
"""Generate ``exp2_table.h``: the table and reduction constants of the
table-driven exp(-x) used by the real-space kernels (se2p_exp_neg).

Method (Tang, "Table-driven implementation of the exponential function in
IEEE floating-point arithmetic", ACM TOMS 15, 1989): with N = 256,
k = rint(-x N / ln 2) <= 0 and r = x + k ln2/N, so |r| <= ln2/(2N) =
1.35e-3, and, with k = N floor(k/N) + j, 0 <= j < N,

    exp(-x) = 2^floor(k/N) * 2^(j/N) * exp(-r).

The table holds 2^(j/N), j = 0..N-1, correctly rounded to double; ln2/N is
split into a double and the double nearest its remainder (each product
k * part is formed exactly inside a fused multiply-add).  k is formed as
fma(x, -N/ln2, 1.5 * 2^52), whose low word then holds k.  exp(-r) is the
degree-4 Taylor polynomial, whose truncation |r|^5/5! is below 3.8e-17.
Values are computed with mpmath at 200 bits.

Run ``python gen_exp_table.py`` to rewrite ``exp2_table.h``.
"""
from __future__ import annotations

import os

import mpmath as mp

N = 256


def main() -> None:
    mp.mp.prec = 200
    ln2 = mp.log(2)
    scale = float(N / ln2)
    hi = float(ln2 / N)
    lo = float(ln2 / N - mp.mpf(hi))
    tab = [float(mp.power(2, mp.mpf(j) / N)) for j in range(N)]
    lines = [
        "// exp2_table.h -- generated by gen_exp_table.py; do not edit.",
        "// Table and reduction constants of se2p_exp_neg (se2p_device.h).",
        "#pragma once",
        "",
        f"#define SE2P_EXP2_N {N}",
        f"#define SE2P_EXP2_SCALE {scale!r}   // N / ln 2",
        f"#define SE2P_EXP2_LN2N_HI {hi!r}   // ln 2 / N, leading part",
        f"#define SE2P_EXP2_LN2N_LO {lo!r}   // and the remainder",
        "#define SE2P_EXP2_SHIFTER 6755399441055744.0   // 1.5 * 2^52",
        "// 2^(j/N), j = 0..N-1, correctly rounded",
        "#define SE2P_EXP2_TABLE \\",
        "    { \\",
    ]
    for i in range(0, N, 4):
        row = ", ".join(repr(v) for v in tab[i:i + 4])
        lines.append(f"        {row}{',' if i + 4 < N else ''} \\")
    lines.append("    }")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "exp2_table.h")
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
