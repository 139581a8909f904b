// se2p_device.h -- device/host helpers shared by the spectral Ewald kernels.
//
// Everything here is FP64.  Functions marked __host__ __device__ are also
// compiled into host code so the CPU test suite can check them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "e1_table.h"

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

// 2^(-j/64), j = 0..63 (correctly rounded; mpmath at 200 bits).
#define SE2P_EXP2_TABLE \
    {1.0, 0.9892280131939755, 0.9785720620877001, 0.9680308967461472, \
     0.9576032806985737, 0.9472879907934828, 0.93708381705515, 0.9269895625416927, \
     0.9170040432046712, 0.9071260877501994, 0.8973545375015536, 0.8876882462632606, \
     0.8781260801866497, 0.8686669176368531, 0.859309649061239, 0.8500531768592617, \
     0.8408964152537145, 0.8318382901633682, 0.8228777390769825, 0.8140137109286739, \
     0.8052451659746271, 0.7965710756711335, 0.7879904225539432, 0.7795022001189185, \
     0.7711054127039704, 0.7627990753722692, 0.7545822137967114, 0.7464538641456324, \
     0.7384130729697497, 0.7304588970903235, 0.7225904034885233, 0.714806669195985, \
     0.7071067811865476, 0.6994898362691556, 0.691954940981916, 0.6845012114872953, \
     0.6771277734684463, 0.6698337620266515, 0.6626183215798707, 0.6554806057623822, \
     0.6484197773255048, 0.6414350080393891, 0.6345254785958666, 0.6276903785123455, \
     0.620928906036742, 0.614240268053435, 0.6076236799902345, 0.6010783657263515, \
     0.5946035575013605, 0.5881984958251406, 0.5818624293887887, 0.5755946149764913, \
     0.5693943173783458, 0.5632608093041209, 0.5571933712979462, 0.5511912916539204, \
     0.5452538663326288, 0.5393803988785599, 0.5335702003384118, 0.5278225891802786, \
     0.5221368912137069, 0.5165124395106142, 0.5109485743270583, 0.5054446430258502}

// exp(-x) for finite x >= 0, table driven (Tang's method with 64 table
// entries): k = rint(64 x / ln 2), r = x - k ln2/64 with ln2/64 split into a
// double and its remainder (Cody-Waite, each product exact inside its fused
// multiply-add), so |r| <= ln2/128; then
// exp(-x) = 2^-(k>>6) * 2^-((k&63)/64) * exp(-r), exp(-r) by its degree-5
// Taylor polynomial (truncation |r|^6/6! < 3.5e-17).  A few ulp; the sums
// that use it are compared with libm's exp at 1e-12.  tab64 = SE2P_EXP2_TABLE.
constexpr int SE2P_EXP2_N = 64;
__host__ __device__ inline double se2p_exp_neg(double x, const double *tab64)
{
    const double k = rint(x * 92.332482616893656);     // 64 / ln 2
    double r = fma(-k, 0.010830424696249145, x);        // ln2/64, leading part
    r = fma(-k, 3.6235106466348430e-19, r);             // and the remainder
    // Estrin's scheme: dependency depth 3
    const double r2 = r * r;
    const double a = 1.0 - r, b = fma(r, -1.0 / 6.0, 0.5), c = fma(r, -1.0 / 120.0, 1.0 / 24.0);
    const double p = fma(r2 * r2, c, fma(r2, b, a));
    const int ki = (int)k;
    return ldexp(tab64[ki & 63] * p, -(ki >> 6));
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
