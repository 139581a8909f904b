// se2p_device.h -- device/host helpers shared by the spectral Ewald kernels.
//
// Everything here is FP64.  Functions marked __host__ __device__ are also
// compiled into host code so the CPU test suite can check them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "e1_table.h"
#include "exp2_table.h"

#define SE2P_PI 3.1415926535897932385
#define SE2P_EULER_GAMMA 0.57721566490153286061

// Packed E1 table row: [scale, shift, c0 .. c14, pad] -> 18 doubles so that a
// row starts on a 16-byte boundary.
#define SE2P_E1_ROW 18

// |d|^2 rounded as round(round(dx*dx) + round(dy*dy)), i.e. without a fused
// multiply-add.  The reference computes r^2 this way
// (mex_stokes_slp_real.cpp:157-159, mex_stokes_dlp_real.cpp:120), and the
// comparisons r^2 == 0 and r^2 < cutoff^2 decide which pairs enter the sum,
// so the rounding must match exactly.
__host__ __device__ inline double se2p_sqsum(double dx, double dy)
{
#pragma clang fp contract(off)
    double a = dx * dx;
    double b = dy * dy;
    return a + b;
}

// exp(-x) for finite x >= 0, table driven (Tang's method, N = 256 table
// entries, exp2_table.h).  k = rint(-x N / ln 2) <= 0 comes from one fused
// multiply-add onto 1.5 * 2^52, which leaves k in the low word of the sum
// (no separate rounding or conversion); r = x + k ln2/N with ln2/N split
// into a double and its remainder (each product exact inside its fused
// multiply-add), so |r| <= ln2/(2N); then, with k = N floor(k/N) + j,
// 0 <= j < N,  exp(-x) = 2^floor(k/N) * 2^(j/N) * exp(-r), exp(-r) by its
// degree-4 Taylor polynomial in Horner form (truncation |r|^5/5! < 3.8e-17).
// A few ulp; the sums that use it are compared with libm's exp at 1e-12.
// tab256 = SE2P_EXP2_TABLE, 2^(j/N) (kernels hold it in LDS).
__host__ __device__ inline double se2p_exp_neg(double x, const double *tab256)
{
    const double kd = fma(x, -SE2P_EXP2_SCALE, SE2P_EXP2_SHIFTER);
    const double k = kd - SE2P_EXP2_SHIFTER;
    uint64_t kb;
    __builtin_memcpy(&kb, &kd, sizeof(kb));
    const int ki = (int)(uint32_t)kb;
    double r = fma(k, SE2P_EXP2_LN2N_HI, x);
    r = fma(k, SE2P_EXP2_LN2N_LO, r);
    const double p = fma(fma(fma(fma(r, 1.0 / 24.0, -1.0 / 6.0), r, 0.5), r, -1.0), r, 1.0);
    return ldexp(tab256[ki & (SE2P_EXP2_N - 1)] * p, ki >> 8);
}

// se2p_sqsum that also returns the two squares (same rounding).
__host__ __device__ inline double se2p_sqsum2(double dx, double dy, double *dx2, double *dy2)
{
#pragma clang fp contract(off)
    double a = dx * dx;
    double b = dy * dy;
    *dx2 = a;
    *dy2 = b;
    return a + b;
}

// Piece index of x > 0 in the E1 table (see gen_e1_table.py).
__host__ __device__ inline int se2p_e1_piece(double x)
{
    if (x < 1.0) return x < 0.5 ? 0 : 1;
    uint64_t bits;
    __builtin_memcpy(&bits, &x, sizeof(bits));
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    if (e >= 10) return SE2P_E1_NPIECE - 1;
    return 2 + 4 * e + (int)((bits >> 50) & 3);
}

// E1(x) for x > 0 given emx = exp(-x) and a packed table (global, constant or
// LDS memory).  Pieces 0-1 hold S(x) = (E1 + gamma + ln x)/x, the rest
// g(x) = e^x E1(x).
__host__ __device__ inline double se2p_e1(double x, double emx, const double *tab)
{
    const double *row = tab + SE2P_E1_ROW * se2p_e1_piece(x);
    double s = x * row[0] + row[1];
    double v = row[2 + SE2P_E1_NCOEF - 1];
#pragma unroll
    for (int k = SE2P_E1_NCOEF - 2; k >= 0; --k) v = v * s + row[2 + k];
    if (x < 1.0) return x * v - SE2P_EULER_GAMMA - log(x);
    return emx * v;
}

// Grid anchor of a point for Gaussian gridding; same arithmetic as the
// reference's FindClosestNode (ewald_tools.cpp:86-114), including the 1e-13
// snapping of the offset and the anchor bump after snapping.
__host__ __device__ inline void se2p_anchor(double x, double y, double Lx, double Ly,
                                            double h, int P, int *mx, int *my,
                                            double *px, double *py)
{
#pragma clang fp contract(off)
    const double snap = 1e-13;
    double hx = Lx / 2.0, hy = Ly / 2.0;
    double ox = x - h * floor(x / h);
    double oy = y - h * floor(y / h);
    int ax = (int)(floor((x + hx) / h) - P / 2);
    int ay = (int)(floor((y + hy) / h) - P / 2);
    if (fabs(ox) < snap || fabs(ox - h) < snap) ox = 0.0;
    if (fabs(oy) < snap || fabs(oy - h) < snap) oy = 0.0;
    if (ox == 0.0 && remainder((x + hx) / h - P / 2, 1.0) < 0.0) ax++;
    if (oy == 0.0 && remainder((y + hy) / h - P / 2, 1.0) < 0.0) ay++;
    *mx = ax;
    *my = ay;
    *px = ox;
    *py = oy;
}

// Cell index used by the real-space cell lists (ewald_tools.cpp:28-35).
__host__ __device__ inline int se2p_cell_of(double x, double y, double Lx, double Ly,
                                            int nsx, int nsy)
{
    int bx = (int)floor(nsx * (x / Lx + 0.5));
    int by = (int)floor(nsy * (y / Ly + 0.5));
    bx = bx < 0 ? 0 : (bx >= nsx ? nsx - 1 : bx);
    by = by < 0 ? 0 : (by >= nsy ? nsy - 1 : by);
    return by * nsx + bx;
}

// Cell-list key of a point: cell index (above) followed by the Morton index
// of its 2^sublog x 2^sublog sub-cell inside the cell (sublog = 0: the cell).
__host__ __device__ inline int se2p_cell_key(double x, double y, double Lx, double Ly, int nsx,
                                             int nsy, int sublog)
{
    int k = se2p_cell_of(x, y, Lx, Ly, nsx, nsy);
    if (sublog > 0) {
        const int sub = 1 << sublog;
        const int bx = k % nsx, by = k / nsx;
        int sx = (int)floor((nsx * (x / Lx + 0.5) - bx) * sub);
        int sy = (int)floor((nsy * (y / Ly + 0.5) - by) * sub);
        sx = sx < 0 ? 0 : (sx >= sub ? sub - 1 : sx);
        sy = sy < 0 ? 0 : (sy >= sub ? sub - 1 : sy);
        int m = 0;
        for (int b = 0; b < sublog; ++b)
            m |= (((sx >> b) & 1) << (2 * b)) | (((sy >> b) & 1) << (2 * b + 1));
        k = (k << (2 * sublog)) | m;
    }
    return k;
}
