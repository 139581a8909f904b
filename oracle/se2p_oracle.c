/*
 * se2p_oracle.c -- CPU restatement of the doubly periodic 2D Stokes spectral
 * Ewald kernels (Stokeslet single layer, stresslet double layer).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  The product (fasttools2d_amd) never links or calls it.
 *
 * It restates, in its own code, the behaviour of the reference MEX kernels
 * (davidkrantz/FastTools2D):
 *   cell lists ............ PeriodicEwald/src/mex/common/ewald_tools.cpp:7-80
 *   grid anchor ........... PeriodicEwald/src/mex/common/ewald_tools.cpp:86-114
 *   SLP real space ........ PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_real.cpp:58-346
 *   SLP Fourier space ..... PeriodicEwald/src/mex/StokesSLP/mex_stokes_slp_kspace.cpp:8-335
 *   DLP real space ........ PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_real.cpp:15-211
 *   DLP Fourier space ..... PeriodicEwald/src/mex/StokesDLP/mex_stokes_dlp_kspace.cpp:10-406
 * The truncation rule, the anchor snapping, the wavenumber/Nyquist convention,
 * the Re/Im extraction and every scaling follow those lines.  Two pieces are
 * written from the mathematics instead of following the reference:
 *   - E1 (exponential integral): power series (Abramowitz & Stegun 5.1.11) for
 *     x <= 1, continued fraction (A&S 5.1.22, evaluated with the modified
 *     Lentz method) for x > 1.  The reference uses tabulated rational
 *     approximations; both are accurate to a few ulp.
 *   - The 2D DFT that the reference borrows from MATLAB's fft2/ifft2: a
 *     radix-2 FFT for power-of-two lengths and Bluestein's algorithm otherwise.
 * Multithreading uses OpenMP exactly where the reference does (cells, sources
 * with per-column locks, k-columns, targets).
 *
 * Pinning (tests/test_oracle_*.py): E1 against mpmath; anchors against brute
 * force; real-space sums against the closed-form image sums restated from
 * PeriodicEwald/tests/matlab_sums/stokes_{slp,dlp}_real_ds.m; Fourier sums
 * against the direct Fourier sums restated from
 * PeriodicEwald/tests/matlab_sums/stokes_{slp,dlp}_kspace_ds.m; the FFT against
 * numpy.fft; the full DLP against the stresslet identity
 * (PeriodicEwald/tests/stresslet_identity_test.m).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.1415926535897932385
#define OR_EULER_GAMMA 0.57721566490153286061

/* ------------------------------------------------------------------------ */
/* Exponential integral E1(x), x > 0.                                        */
/* ------------------------------------------------------------------------ */
double or_e1(double x)
{
    if (!(x > 0.0)) return INFINITY;
    if (x <= 1.0) {
        /* E1(x) = -gamma - ln x - sum_{k>=1} (-x)^k / (k k!)   (A&S 5.1.11) */
        double term = 1.0;   /* (-x)^k / k! */
        double sum = 0.0;
        for (int k = 1; k < 60; ++k) {
            term *= -x / (double)k;
            double add = term / (double)k;
            sum += add;
            if (fabs(add) < 1e-18 * fabs(sum)) break;
        }
        return -OR_EULER_GAMMA - log(x) - sum;
    }
    /* E1(x) = e^{-x} / (x+1 - 1/(x+3 - 4/(x+5 - 9/(x+7 - ...))))  (A&S 5.1.22),
     * modified Lentz evaluation. */
    const double tiny = 1e-300;
    double b = x + 1.0;
    double c = 1.0 / tiny;
    double d = 1.0 / b;
    double h = d;
    for (int i = 1; i < 10000; ++i) {
        double an = -(double)i * (double)i;
        b += 2.0;
        d = 1.0 / (an * d + b);
        c = b + an / c;
        double del = c * d;
        h *= del;
        if (fabs(del - 1.0) < 1e-17) break;
    }
    return h * exp(-x);
}

/* ------------------------------------------------------------------------ */
/* Cell lists: stable counting sort of points into nsx*nsy boxes.            */
/* Follows ewald_tools.cpp:7-80 (box index with clamping :28-35, prefix sum   */
/* :57-68, stable scatter :71-74).                                           */
/* ------------------------------------------------------------------------ */
static int or_box_of(double x, double y, double Lx, double Ly, int nsx, int nsy)
{
    int bx = (int)floor(nsx * (x / Lx + 0.5));
    int by = (int)floor(nsy * (y / Ly + 0.5));
    if (bx < 0) bx = 0;
    if (bx >= nsx) bx = nsx - 1;
    if (by < 0) by = 0;
    if (by >= nsy) by = nsy - 1;
    return by * nsx + bx;
}

/* perm[k] = original index of the k-th point in box order; off[0..nbox]. */
void or_cell_lists(const double *p, int64_t n, double Lx, double Ly,
                   int nsx, int nsy, int64_t *perm, int64_t *off)
{
    int nbox = nsx * nsy;
    int *box = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nbox + 1));
    for (int b = 0; b <= nbox; ++b) off[b] = 0;
    for (int64_t i = 0; i < n; ++i) {
        box[i] = or_box_of(p[2 * i], p[2 * i + 1], Lx, Ly, nsx, nsy);
        off[box[i] + 1]++;
    }
    for (int b = 0; b < nbox; ++b) off[b + 1] += off[b];
    for (int b = 0; b <= nbox; ++b) fill[b] = off[b];
    for (int64_t i = 0; i < n; ++i) perm[fill[box[i]]++] = i;
    free(box);
    free(fill);
}

/* ------------------------------------------------------------------------ */
/* Grid anchor of a point for Gaussian gridding (ewald_tools.cpp:86-114).    */
/* mx,my: first grid node of the (P+1)^2 support; px,py: offset of the point  */
/* from the grid line below it, snapped to 0 within 1e-13 of a node.          */
/* ------------------------------------------------------------------------ */
void or_find_closest_node(double x, double y, double Lx, double Ly, double h,
                          int P, int *mx, int *my, double *px, double *py)
{
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

/* ------------------------------------------------------------------------ */
/* Real-space sums.                                                          */
/* ------------------------------------------------------------------------ */
static const int OR_NBX[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
static const int OR_NBY[8] = {-1, 0, 1, -1, 1, -1, 0, 1};

/* Stokeslet real-space pair term, accumulated into (a1,a2), before the 1/4pi.
 * d = difference vector (either sign), r2 = |d|^2 > 0.
 * 0.5 E1(xi^2 r^2) f + e^{-xi^2 r^2} ((d.f) d / r^2 - f)
 * (mex_stokes_slp_real.cpp:166-225, closed form stokes_slp_real_ds.m:470-477) */
static inline void or_slp_pair(double dx, double dy, double r2, double xi,
                               double f1, double f2, double *a1, double *a2)
{
    double x = r2 * xi * xi;
    double e = exp(-x);
    double E = or_e1(x);
    double c = dx * f2 - dy * f1;
    double g = e / r2;
    *a1 += 0.5 * E * f1 + g * dy * c;
    *a2 += 0.5 * E * f2 - g * dx * c;
}

int or_slp_real(const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                const double *f, double xi, int nsx, int nsy, double Lx,
                double Ly, double *u)
{
    int nbox = nsx * nsy;
    int64_t *ps = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ns + 1));
    int64_t *pt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nt + 1));
    int64_t *os = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nbox + 1));
    int64_t *ot = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nbox + 1));
    or_cell_lists(psrc, ns, Lx, Ly, nsx, nsy, ps, os);
    or_cell_lists(ptar, nt, Lx, Ly, nsx, nsy, pt, ot);
    double *xs = (double *)malloc(sizeof(double) * (size_t)(2 * ns + 2));
    double *fs = (double *)malloc(sizeof(double) * (size_t)(2 * ns + 2));
    double *xt = (double *)malloc(sizeof(double) * (size_t)(2 * nt + 2));
    double *acc = (double *)calloc((size_t)(2 * nt + 2), sizeof(double));
    for (int64_t k = 0; k < ns; ++k) {
        xs[2 * k] = psrc[2 * ps[k]];
        xs[2 * k + 1] = psrc[2 * ps[k] + 1];
        fs[2 * k] = f[2 * ps[k]];
        fs[2 * k + 1] = f[2 * ps[k] + 1];
    }
    for (int64_t k = 0; k < nt; ++k) {
        xt[2 * k] = ptar[2 * pt[k]];
        xt[2 * k + 1] = ptar[2 * pt[k] + 1];
    }
    /* coincident-point limit: -(gamma/2 + 1 + ln xi) (mex_stokes_slp_real.cpp:124) */
    const double self = -(0.5 * OR_EULER_GAMMA + 1.0) - log(xi);
    const double cut2 = Lx * Ly / nsx / nsy;   /* mex_stokes_slp_real.cpp:127 */

#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbox; ++b) {
        if (ot[b + 1] == ot[b]) continue;
        int bx = b % nsx, by = b / nsx;
        for (int64_t t = ot[b]; t < ot[b + 1]; ++t) {
            double tx = xt[2 * t], ty = xt[2 * t + 1];
            double a1 = 0.0, a2 = 0.0;
            /* own box: every pair, no cutoff (:150-230) */
            for (int64_t s = os[b]; s < os[b + 1]; ++s) {
                double dx = tx - xs[2 * s];
                double dy = ty - xs[2 * s + 1];
                double r2 = dx * dx + dy * dy;
                if (r2 == 0.0) {
                    a1 += self * fs[2 * s];
                    a2 += self * fs[2 * s + 1];
                } else {
                    or_slp_pair(dx, dy, r2, xi, fs[2 * s], fs[2 * s + 1], &a1, &a2);
                }
            }
            /* eight neighbour boxes with periodic offset, r^2 < cut2 (:233-332) */
            for (int j = 0; j < 8; ++j) {
                int qx = bx + OR_NBX[j], qy = by + OR_NBY[j];
                int wx = (qx + nsx) % nsx, wy = (qy + nsy) % nsy;
                int sb = wy * nsx + wx;
                if (os[sb + 1] == os[sb]) continue;
                double zx = (Lx * (qx - wx)) / nsx;
                double zy = (Ly * (qy - wy)) / nsy;
                double sx = tx - zx, sy = ty - zy;
                for (int64_t s = os[sb]; s < os[sb + 1]; ++s) {
                    double dx = xs[2 * s] - sx;
                    double dy = xs[2 * s + 1] - sy;
                    double r2 = dx * dx + dy * dy;
                    if (r2 < cut2)
                        or_slp_pair(dx, dy, r2, xi, fs[2 * s], fs[2 * s + 1], &a1, &a2);
                }
            }
            acc[2 * t] = a1;
            acc[2 * t + 1] = a2;
        }
    }
    for (int64_t k = 0; k < nt; ++k) {
        u[2 * pt[k]] = acc[2 * k] / (4 * OR_PI);
        u[2 * pt[k] + 1] = acc[2 * k + 1] / (4 * OR_PI);
    }
    free(ps); free(pt); free(os); free(ot);
    free(xs); free(fs); free(xt); free(acc);
    return 0;
}

/* Stresslet real-space pair term (mex_stokes_dlp_real.cpp:125-138), before 1/4pi.
 * (d1,d2) = source - target (+ image offset).  q = (S11, S12+S21, S22). */
static inline void or_dlp_pair(double d1, double d2, double r2, double xi2,
                               double q11, double q12, double q22,
                               double *a1, double *a2)
{
    double e2 = exp(-xi2 * r2);
    double pre = 2 * xi2;
    double fb = -4 * (1 + xi2 * r2) / r2 / r2;
    double t111 = d1 * d1 * d1 * fb + pre * 3 * d1;
    double t112 = d1 * d1 * d2 * fb + pre * d2;
    double t122 = d1 * d2 * d2 * fb + pre * d1;
    double t222 = d2 * d2 * d2 * fb + pre * 3 * d2;
    /* T211 = T112 and T212 = T122 by symmetry of the stresslet tensor */
    *a1 += e2 * (t111 * q11 + t112 * q12 + t122 * q22);
    *a2 += e2 * (t112 * q11 + t122 * q12 + t222 * q22);
}

int or_dlp_real(const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                const double *f, const double *nv, double xi, int nsx, int nsy,
                double Lx, double Ly, double *u)
{
    int nbox = nsx * nsy;
    int64_t *ps = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ns + 1));
    int64_t *pt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nt + 1));
    int64_t *os = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nbox + 1));
    int64_t *ot = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nbox + 1));
    or_cell_lists(psrc, ns, Lx, Ly, nsx, nsy, ps, os);
    or_cell_lists(ptar, nt, Lx, Ly, nsx, nsy, pt, ot);
    double *xs = (double *)malloc(sizeof(double) * (size_t)(2 * ns + 2));
    double *qs = (double *)malloc(sizeof(double) * (size_t)(3 * ns + 3));
    double *xt = (double *)malloc(sizeof(double) * (size_t)(2 * nt + 2));
    double *acc = (double *)calloc((size_t)(2 * nt + 2), sizeof(double));
    for (int64_t k = 0; k < ns; ++k) {
        int64_t i = ps[k];
        xs[2 * k] = psrc[2 * i];
        xs[2 * k + 1] = psrc[2 * i + 1];
        /* S = f (x) n (mex_stokes_dlp_real.cpp:86-89); S12 and S21 only ever
         * appear summed (:133,138). */
        qs[3 * k] = f[2 * i] * nv[2 * i];
        qs[3 * k + 1] = f[2 * i] * nv[2 * i + 1] + f[2 * i + 1] * nv[2 * i];
        qs[3 * k + 2] = f[2 * i + 1] * nv[2 * i + 1];
    }
    for (int64_t k = 0; k < nt; ++k) {
        xt[2 * k] = ptar[2 * pt[k]];
        xt[2 * k + 1] = ptar[2 * pt[k] + 1];
    }
    const double cut2 = Lx * Ly / nsx / nsy;   /* :99 */
    const double xi2 = xi * xi;

#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbox; ++b) {
        if (ot[b + 1] == ot[b]) continue;
        int bx = b % nsx, by = b / nsx;
        for (int64_t t = ot[b]; t < ot[b + 1]; ++t) {
            double tx = xt[2 * t], ty = xt[2 * t + 1];
            double a1 = 0.0, a2 = 0.0;
            for (int64_t s = os[b]; s < os[b + 1]; ++s) {   /* :113-140 */
                double d1 = xs[2 * s] - tx;
                double d2 = xs[2 * s + 1] - ty;
                double r2 = d1 * d1 + d2 * d2;
                if (r2 == 0.0) continue;
                or_dlp_pair(d1, d2, r2, xi2, qs[3 * s], qs[3 * s + 1], qs[3 * s + 2], &a1, &a2);
            }
            for (int j = 0; j < 8; ++j) {                    /* :143-186 */
                int qx = bx + OR_NBX[j], qy = by + OR_NBY[j];
                int wx = (qx + nsx) % nsx, wy = (qy + nsy) % nsy;
                int sb = wy * nsx + wx;
                if (os[sb + 1] == os[sb]) continue;
                double zx = (Lx * (qx - wx)) / nsx;
                double zy = (Ly * (qy - wy)) / nsy;
                for (int64_t s = os[sb]; s < os[sb + 1]; ++s) {
                    double d1 = xs[2 * s] - tx + zx;
                    double d2 = xs[2 * s + 1] - ty + zy;
                    double r2 = d1 * d1 + d2 * d2;
                    if (r2 < cut2)
                        or_dlp_pair(d1, d2, r2, xi2, qs[3 * s], qs[3 * s + 1], qs[3 * s + 2], &a1, &a2);
                }
            }
            acc[2 * t] = a1;
            acc[2 * t + 1] = a2;
        }
    }
    for (int64_t k = 0; k < nt; ++k) {
        u[2 * pt[k]] = acc[2 * k] / 4 / OR_PI;
        u[2 * pt[k] + 1] = acc[2 * k + 1] / 4 / OR_PI;
    }
    free(ps); free(pt); free(os); free(ot);
    free(xs); free(qs); free(xt); free(acc);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 1D / 2D complex DFT (stand-in for MATLAB fft2/ifft2).                     */
/* Forward: X_k = sum_n x_n e^{-2 pi i k n / N}; inverse has e^{+...} / N.   */
/* ------------------------------------------------------------------------ */
static int or_is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* in-place iterative radix-2, sign = -1 forward, +1 backward (unscaled) */
static void or_fft_pow2(double *re, double *im, int n, int sign)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            double ang = sign * 2.0 * OR_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (int i = k; i < n; i += len) {
                int j = i + half;
                double xr = re[j] * wr - im[j] * wi;
                double xi = re[j] * wi + im[j] * wr;
                re[j] = re[i] - xr; im[j] = im[i] - xi;
                re[i] += xr;        im[i] += xi;
            }
        }
    }
}

/* Bluestein (chirp-z) for arbitrary n via a power-of-two convolution. */
static void or_fft_any(double *re, double *im, int n, int sign)
{
    if (n <= 1) return;
    if (or_is_pow2(n)) { or_fft_pow2(re, im, n, sign); return; }
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    double *ar = (double *)calloc((size_t)m, sizeof(double));
    double *ai = (double *)calloc((size_t)m, sizeof(double));
    double *br = (double *)calloc((size_t)m, sizeof(double));
    double *bi = (double *)calloc((size_t)m, sizeof(double));
    double *cr = (double *)calloc((size_t)n, sizeof(double));
    double *ci = (double *)calloc((size_t)n, sizeof(double));
    for (int k = 0; k < n; ++k) {
        /* chirp w_k = e^{sign i pi k^2 / n}; k^2 mod 2n keeps the angle small */
        long long k2 = ((long long)k * k) % (2LL * n);
        double ang = sign * OR_PI * (double)k2 / (double)n;
        cr[k] = cos(ang);
        ci[k] = sin(ang);
    }
    for (int k = 0; k < n; ++k) {
        ar[k] = re[k] * cr[k] - im[k] * ci[k];
        ai[k] = re[k] * ci[k] + im[k] * cr[k];
    }
    br[0] = cr[0]; bi[0] = -ci[0];
    for (int k = 1; k < n; ++k) {
        br[k] = cr[k]; bi[k] = -ci[k];
        br[m - k] = cr[k]; bi[m - k] = -ci[k];
    }
    or_fft_pow2(ar, ai, m, -1);
    or_fft_pow2(br, bi, m, -1);
    for (int k = 0; k < m; ++k) {
        double xr = ar[k] * br[k] - ai[k] * bi[k];
        double xi = ar[k] * bi[k] + ai[k] * br[k];
        ar[k] = xr; ai[k] = xi;
    }
    or_fft_pow2(ar, ai, m, +1);
    for (int k = 0; k < n; ++k) {
        double xr = ar[k] / m, xi = ai[k] / m;
        re[k] = xr * cr[k] - xi * ci[k];
        im[k] = xr * ci[k] + xi * cr[k];
    }
    free(ar); free(ai); free(br); free(bi); free(cr); free(ci);
}

/* 2D transform of an My x Mx column-major array (index x*My + y). */
void or_fft2(double *re, double *im, int Mx, int My, int sign)
{
#pragma omp parallel
    {
        double *tr = (double *)malloc(sizeof(double) * (size_t)(Mx > My ? Mx : My));
        double *ti = (double *)malloc(sizeof(double) * (size_t)(Mx > My ? Mx : My));
#pragma omp for
        for (int x = 0; x < Mx; ++x) {
            memcpy(tr, re + (size_t)x * My, sizeof(double) * (size_t)My);
            memcpy(ti, im + (size_t)x * My, sizeof(double) * (size_t)My);
            or_fft_any(tr, ti, My, sign);
            memcpy(re + (size_t)x * My, tr, sizeof(double) * (size_t)My);
            memcpy(im + (size_t)x * My, ti, sizeof(double) * (size_t)My);
        }
#pragma omp for
        for (int y = 0; y < My; ++y) {
            for (int x = 0; x < Mx; ++x) {
                tr[x] = re[(size_t)x * My + y];
                ti[x] = im[(size_t)x * My + y];
            }
            or_fft_any(tr, ti, Mx, sign);
            for (int x = 0; x < Mx; ++x) {
                re[(size_t)x * My + y] = tr[x];
                im[(size_t)x * My + y] = ti[x];
            }
        }
        free(tr);
        free(ti);
    }
    if (sign > 0) {
        double s = 1.0 / ((double)Mx * (double)My);
        size_t g = (size_t)Mx * My;
        for (size_t i = 0; i < g; ++i) { re[i] *= s; im[i] *= s; }
    }
}

/* ------------------------------------------------------------------------ */
/* Fast Gaussian gridding: spreading and interpolation.                      */
/* Weight of node (mx+i, my+j) for a point with anchor (mx,my), offset       */
/* (px,py):  exp(-2 xi^2/eta * |node - point|^2), factored as in             */
/* mex_stokes_slp_kspace.cpp:64-67,95-99 into a shared 1D table e1[] and four */
/* exponentials per point.                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int mx, my;
    double ex, e4y, e3x, e3y;
} or_fgg_point;

static void or_fgg_setup(double x, double y, double Lx, double Ly, double h,
                         int P, double xi, double eta, double w,
                         or_fgg_point *g)
{
    double px, py;
    or_find_closest_node(x, y, Lx, Ly, h, P, &g->mx, &g->my, &px, &py);
    double c = -2 * xi * xi / eta;
    g->ex = exp(c * (px * px + py * py + 2 * w * px));
    g->e4y = exp(2 * c * w * py);
    g->e3x = exp(-2 * c * h * px);
    g->e3y = exp(-2 * c * h * py);
}

/* nq densities per source are spread to nq grids (each My x Mx). */
static void or_spread(const double *psrc, int64_t ns, const double *q, int nq,
                      double xi, double eta, int Mx, int My, double Lx,
                      double Ly, double w, int P, const double *e1, double **H)
{
    double h = Lx / Mx;
#ifdef _OPENMP
    omp_lock_t *locks = (omp_lock_t *)malloc(sizeof(omp_lock_t) * (size_t)Mx);
    for (int j = 0; j < Mx; ++j) omp_init_lock(&locks[j]);
#endif
#pragma omp parallel for
    for (int64_t k = 0; k < ns; ++k) {
        or_fgg_point g;
        or_fgg_setup(psrc[2 * k], psrc[2 * k + 1], Lx, Ly, h, P, xi, eta, w, &g);
        double ex = g.ex;
        for (int i = 0; i <= P; ++i) {
            int col = ((i + g.mx) % Mx + Mx) % Mx;
            double ey = ex * g.e4y * e1[i];
#ifdef _OPENMP
            omp_set_lock(&locks[col]);
#endif
            for (int j = 0; j <= P; ++j) {
                int row = ((j + g.my) % My + My) % My;
                double wgt = ey * e1[j];
                size_t idx = (size_t)col * My + row;
                for (int c = 0; c < nq; ++c) H[c][idx] += wgt * q[(size_t)nq * k + c];
                ey *= g.e3y;
            }
#ifdef _OPENMP
            omp_unset_lock(&locks[col]);
#endif
            ex *= g.e3x;
        }
    }
#ifdef _OPENMP
    for (int j = 0; j < Mx; ++j) omp_destroy_lock(&locks[j]);
    free(locks);
#endif
}

static void or_interp(const double *ptar, int64_t nt, double xi, double eta,
                      int Mx, int My, double Lx, double Ly, double w, int P,
                      const double *e1, const double *G1, const double *G2,
                      double *u)
{
    double h = Lx / Mx;
    /* (4 xi^2/eta)^2 h^2 / pi / (4 pi)  (mex_stokes_slp_kspace.cpp:325-328) */
    double s = 4 * xi * xi / eta;
    s = s * s * h * h / OR_PI;
#pragma omp parallel for
    for (int64_t k = 0; k < nt; ++k) {
        or_fgg_point g;
        or_fgg_setup(ptar[2 * k], ptar[2 * k + 1], Lx, Ly, h, P, xi, eta, w, &g);
        double ex = g.ex, a1 = 0.0, a2 = 0.0;
        for (int i = 0; i <= P; ++i) {
            int col = ((i + g.mx) % Mx + Mx) % Mx;
            double ey = ex * g.e4y * e1[i];
            for (int j = 0; j <= P; ++j) {
                int row = ((j + g.my) % My + My) % My;
                double wgt = ey * e1[j];
                size_t idx = (size_t)col * My + row;
                a1 += wgt * G1[idx];
                a2 += wgt * G2[idx];
                ey *= g.e3y;
            }
            ex *= g.e3x;
        }
        u[2 * k] = a1 * (s / (4 * OR_PI));
        u[2 * k + 1] = a2 * (s / (4 * OR_PI));
    }
}

/* Wavenumber of FFT index j on a length-M axis (mex_stokes_slp_kspace.cpp:193-198,
 * 210-211): j <= M/2 maps to +j, the rest to j - M.  The Nyquist index is +. */
static inline double or_wavenumber(int j, int M, double L)
{
    return (j <= M / 2) ? 2.0 * OR_PI / L * j : 2.0 * OR_PI / L * (j - M);
}

static double *or_e1_table(double xi, double eta, double h, int P)
{
    double *e1 = (double *)malloc(sizeof(double) * (size_t)(P + 1));
    double c = -2 * xi * xi / eta * h * h;
    for (int j = -P / 2; j <= P / 2; ++j) e1[j + P / 2] = exp(c * j * j);
    return e1;
}

/* Filtered Stokeslet grids: spread f at the points, forward DFT, Stokeslet
 * filter, k = 0 mode zeroed, inverse DFT; G1, G2 (Mx*My each, caller-owned,
 * zeroed on entry) receive the real parts (mex_stokes_slp_kspace.cpp:59-240). */
static void or_slp_kspace_grids(const double *p, int64_t n, double xi, double eta,
                                const double *f, int Mx, int My, double Lx, double Ly,
                                double w, int P, const double *e1, double *G1, double *G2)
{
    size_t G = (size_t)Mx * My;
    double *H1i = (double *)calloc(G, sizeof(double));
    double *H2i = (double *)calloc(G, sizeof(double));
    double *H1r = G1, *H2r = G2;
    double *H[2] = {H1r, H2r};
    or_spread(p, n, f, 2, xi, eta, Mx, My, Lx, Ly, w, P, e1, H);
    or_fft2(H1r, H1i, Mx, My, -1);
    or_fft2(H2r, H2i, Mx, My, -1);
    /* Stokeslet filter (mex_stokes_slp_kspace.cpp:189-223) */
#pragma omp parallel for
    for (int j = 0; j < Mx; ++j) {
        double k1 = or_wavenumber(j, Mx, Lx);
        for (int k = 0; k < My; ++k) {
            double k2 = or_wavenumber(k, My, Ly);
            double K2 = k1 * k1 + k2 * k2;
            double e = (1.0 / (K2 * K2) + 0.25 / (K2 * xi * xi)) *
                       exp(-0.25 * (1 - eta) / (xi * xi) * K2);
            size_t p = (size_t)j * My + k;
            double t = k2 * (k2 * H1r[p] - k1 * H2r[p]);
            H2r[p] = k1 * (k1 * H2r[p] - k2 * H1r[p]) * e;
            H1r[p] = t * e;
            t = k2 * (k2 * H1i[p] - k1 * H2i[p]);
            H2i[p] = k1 * (k1 * H2i[p] - k2 * H1i[p]) * e;
            H1i[p] = t * e;
        }
    }
    H1r[0] = H2r[0] = H1i[0] = H2i[0] = 0.0;   /* :226-229 */
    or_fft2(H1r, H1i, Mx, My, +1);
    or_fft2(H2r, H2i, Mx, My, +1);
    free(H1i);
    free(H2i);
}

int or_slp_kspace(const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                  double xi, double eta, const double *f, int Mx, int My,
                  double Lx, double Ly, double w, int P, double *u)
{
    size_t G = (size_t)Mx * My;
    double h = Lx / Mx;
    double *e1 = or_e1_table(xi, eta, h, P);
    double *G1 = (double *)calloc(G, sizeof(double));
    double *G2 = (double *)calloc(G, sizeof(double));
    or_slp_kspace_grids(psrc, ns, xi, eta, f, Mx, My, Lx, Ly, w, P, e1, G1, G2);
    /* real part (:240) */
    or_interp(ptar, nt, xi, eta, Mx, My, Lx, Ly, w, P, e1, G1, G2, u);
    free(e1); free(G1); free(G2);
    return 0;
}

/* Stresslet filter coefficients at wavenumber (k1, k2) (mex_stokes_dlp_kspace.cpp:
 * 232-247): et = e*t and b111, b112, b122, b222. */
static void or_dlp_filter_coef(double k1, double k2, double xi, double eta, double *et,
                               double *b111, double *b112, double *b122, double *b222)
{
    double K2 = k1 * k1 + k2 * k2;
    double e = exp(-0.25 * (1 - eta) / (xi * xi) * K2) / K2 / K2;
    double t = 1 + 0.25 * K2 / xi / xi;
    *et = e * t;
    *b111 = 3.0 * k1 * K2 - 2.0 * k1 * k1 * k1;
    *b112 = k2 * K2 - 2.0 * k1 * k1 * k2;
    *b122 = k1 * K2 - 2.0 * k1 * k2 * k2;
    *b222 = 3.0 * k2 * K2 - 2.0 * k2 * k2 * k2;
}

/* The four stresslet densities n_l f_j in the reference's grid order
 * (mex_stokes_dlp_kspace.cpp:129-132): (n1 f1, n2 f1, n1 f2, n2 f2). */
static double *or_dlp_q4(const double *f, const double *nv, int64_t n)
{
    double *q = (double *)malloc(sizeof(double) * (size_t)(4 * n + 4));
    for (int64_t k = 0; k < n; ++k) {
        q[4 * k] = nv[2 * k] * f[2 * k];
        q[4 * k + 1] = nv[2 * k + 1] * f[2 * k];
        q[4 * k + 2] = nv[2 * k] * f[2 * k + 1];
        q[4 * k + 3] = nv[2 * k + 1] * f[2 * k + 1];
    }
    return q;
}

/* Filtered stresslet grids: imaginary parts of the two inverse transforms
 * (mex_stokes_dlp_kspace.cpp:129-305) into G1, G2 (caller-owned). */
static void or_dlp_kspace_grids(const double *psrc, int64_t ns, double xi, double eta,
                                const double *f, const double *nv, int Mx, int My,
                                double Lx, double Ly, double w, int P, const double *e1,
                                double *G1, double *G2)
{
    size_t G = (size_t)Mx * My;
    double *Hr[4], *Hi[4];
    for (int c = 0; c < 4; ++c) {
        Hr[c] = (double *)calloc(G, sizeof(double));
        Hi[c] = (double *)calloc(G, sizeof(double));
    }
    double *q = or_dlp_q4(f, nv, ns);
    or_spread(psrc, ns, q, 4, xi, eta, Mx, My, Lx, Ly, w, P, e1, Hr);
    for (int c = 0; c < 4; ++c) or_fft2(Hr[c], Hi[c], Mx, My, -1);
    /* stresslet filter (:219-279) */
#pragma omp parallel for
    for (int j = 0; j < Mx; ++j) {
        double k1 = or_wavenumber(j, Mx, Lx);
        for (int k = 0; k < My; ++k) {
            double k2 = or_wavenumber(k, My, Ly);
            double et, b111, b112, b122, b222;
            or_dlp_filter_coef(k1, k2, xi, eta, &et, &b111, &b112, &b122, &b222);
            size_t p = (size_t)j * My + k;
            double m1r = Hr[1][p] + Hr[2][p], m1i = Hi[1][p] + Hi[2][p];
            double o1r = et * (b111 * Hr[0][p] + b112 * m1r + b122 * Hr[3][p]);
            double o1i = et * (b111 * Hi[0][p] + b112 * m1i + b122 * Hi[3][p]);
            double o2r = et * (b112 * Hr[0][p] + b122 * m1r + b222 * Hr[3][p]);
            double o2i = et * (b112 * Hi[0][p] + b122 * m1i + b222 * Hi[3][p]);
            Hr[0][p] = o1r; Hi[0][p] = o1i;
            Hr[1][p] = o2r; Hi[1][p] = o2i;
        }
    }
    Hr[0][0] = Hr[1][0] = Hi[0][0] = Hi[1][0] = 0.0;   /* :282-285 */
    or_fft2(Hr[0], Hi[0], Mx, My, +1);
    or_fft2(Hr[1], Hi[1], Mx, My, +1);
    /* imaginary part (:304-305) */
    memcpy(G1, Hi[0], sizeof(double) * G);
    memcpy(G2, Hi[1], sizeof(double) * G);
    for (int c = 0; c < 4; ++c) { free(Hr[c]); free(Hi[c]); }
    free(q);
}

int or_dlp_kspace(const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                  double xi, double eta, const double *f, const double *nv,
                  int Mx, int My, double Lx, double Ly, double w, int P,
                  double *u)
{
    size_t G = (size_t)Mx * My;
    double h = Lx / Mx;
    double *e1 = or_e1_table(xi, eta, h, P);
    double *G1 = (double *)malloc(sizeof(double) * G);
    double *G2 = (double *)malloc(sizeof(double) * G);
    or_dlp_kspace_grids(psrc, ns, xi, eta, f, nv, Mx, My, Lx, Ly, w, P, e1, G1, G2);
    or_interp(ptar, nt, xi, eta, Mx, My, Lx, Ly, w, P, e1, G1, G2, u);
    free(G1); free(G2);
    free(e1);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Adjoints (gradients) of the four sums.                                    */
/* With u the output of a sum and L = sum_t du_t . u_t, each *_adj function  */
/* returns dL/dpsrc, dL/dptar, dL/df (and dL/dn for the stresslet); NULL     */
/* outputs are skipped.  The reference has no adjoint: these are derived     */
/* here from the forward restatements above.  Pair inclusion, image offsets  */
/* and the coincident-point rules are those of the forward sums; gradients   */
/* treat the cell and anchor assignment as fixed (piecewise constant), and   */
/* a coincident pair (r^2 == 0, a constant term) contributes no position     */
/* gradient.  Written with full tensor derivatives (index loops), not the    */
/* contracted forms of the product, so that the two are independent.        */
/* ------------------------------------------------------------------------ */

/* Stokeslet pair G(d) = 0.5 E1(xi^2 r^2) I + e^{-xi^2 r^2} (d d^T / r^2 - I)
 * and its derivative dG_ij/dd_l.  v += G a;  g_l += sum_ij a_i b_j dG_ij/dd_l. */
static void or_slp_pair_adj(const double d[2], double r2, double xi, const double a[2],
                            const double b[2], double v[2], double g[2])
{
    double xi2 = xi * xi;
    double e = exp(-xi2 * r2);
    double E = or_e1(xi2 * r2);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double dij = (i == j) ? 1.0 : 0.0;
            double Gij = 0.5 * E * dij + e * (d[i] * d[j] / r2 - dij);
            v[i] += Gij * a[j];
            for (int l = 0; l < 2; ++l) {
                double dil = (i == l) ? 1.0 : 0.0, djl = (j == l) ? 1.0 : 0.0;
                /* d/dd_l of 0.5 E1(xi^2 r^2) = -e d_l / r^2 (dE1/dx = -e^{-x}/x) */
                double dG = -e * d[l] / r2 * dij
                            - 2.0 * xi2 * e * d[l] * (d[i] * d[j] / r2 - dij)
                            + e * ((dil * d[j] + djl * d[i]) / r2 -
                                   2.0 * d[i] * d[j] * d[l] / (r2 * r2));
                g[l] += a[i] * b[j] * dG;
            }
        }
}

/* Stresslet pair T_ijk(d) = e2 (fb d_i d_j d_k + 2 xi^2 (δij dk + δik dj + δjk di)),
 * fb = -4 (1 + xi^2 r^2) / r^4, e2 = e^{-xi^2 r^2} (mex_stokes_dlp_real.cpp:125-138),
 * contracted with a (the output weight) and S = f (x) n:
 *   M_jk += sum_i a_i T_ijk;   g_l += sum_ijk a_i S_jk dT_ijk/dd_l. */
static void or_dlp_pair_adj(const double d[2], double r2, double xi, const double a[2],
                            const double S[2][2], double M[2][2], double g[2])
{
    double xi2 = xi * xi;
    double e2 = exp(-xi2 * r2);
    double fb = -4.0 * (1.0 + xi2 * r2) / (r2 * r2);
    double fbp = 4.0 * (2.0 + xi2 * r2) / (r2 * r2 * r2);   /* d fb / d(r^2) */
    double pre = 2.0 * xi2;
#define OR_KD(p, q) ((p) == (q) ? 1.0 : 0.0)
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                double sym = OR_KD(i, j) * d[k] + OR_KD(i, k) * d[j] + OR_KD(j, k) * d[i];
                double T = e2 * (fb * d[i] * d[j] * d[k] + pre * sym);
                M[j][k] += a[i] * T;
                for (int l = 0; l < 2; ++l) {
                    double dsym = OR_KD(i, j) * OR_KD(k, l) + OR_KD(i, k) * OR_KD(j, l) +
                                  OR_KD(j, k) * OR_KD(i, l);
                    double dT = -2.0 * xi2 * d[l] * T +
                                e2 * (2.0 * fbp * d[l] * d[i] * d[j] * d[k] +
                                      fb * (OR_KD(i, l) * d[j] * d[k] + OR_KD(j, l) * d[i] * d[k] +
                                            OR_KD(k, l) * d[i] * d[j]) +
                                      pre * dsym);
                    g[l] += a[i] * S[j][k] * dT;
                }
            }
#undef OR_KD
}

/* Real-space adjoint, both kernels.  Targets are visited as in the forward
 * sums (cells in parallel); contributions to sources go to per-thread buffers
 * that are added at the end.  Targets with du == 0 are skipped (they
 * contribute nothing). */
static int or_real_adj(int dlp, const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                       const double *f, const double *nv, double xi, int nsx, int nsy,
                       double Lx, double Ly, const double *du, double *dpsrc,
                       double *dptar, double *df, double *dn)
{
    int nbox = nsx * nsy;
    int64_t *ps = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ns + 1));
    int64_t *pt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nt + 1));
    int64_t *os = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nbox + 1));
    int64_t *ot = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nbox + 1));
    or_cell_lists(psrc, ns, Lx, Ly, nsx, nsy, ps, os);
    or_cell_lists(ptar, nt, Lx, Ly, nsx, nsy, pt, ot);
    const double self = -(0.5 * OR_EULER_GAMMA + 1.0) - log(xi);   /* slp :124 */
    const double cut2 = Lx * Ly / nsx / nsy;
    /* per source: (dL/dx, dL/dy, then 2 (SLP: G du) or 4 (DLP: M) entries) */
    const int W = dlp ? 6 : 4;
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    double *sbuf = (double *)calloc((size_t)nthr * (size_t)(W * ns + 1), sizeof(double));
    double *tacc = (double *)calloc((size_t)(2 * nt + 2), sizeof(double));

#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nbox; ++b) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *mine = sbuf + (size_t)tid * (size_t)(W * ns + 1);
        int bx = b % nsx, by = b / nsx;
        for (int64_t tk = ot[b]; tk < ot[b + 1]; ++tk) {
            int64_t t = pt[tk];
            double a[2] = {du[2 * t], du[2 * t + 1]};
            if (a[0] == 0.0 && a[1] == 0.0) continue;
            double tx = ptar[2 * t], ty = ptar[2 * t + 1];
            double gt[2] = {0.0, 0.0};
            for (int j = -1; j < 8; ++j) {   /* -1: own box */
                int sb, own = (j < 0);
                double zx = 0.0, zy = 0.0;
                if (own) {
                    sb = b;
                } else {
                    int qx = bx + OR_NBX[j], qy = by + OR_NBY[j];
                    int wx = (qx + nsx) % nsx, wy = (qy + nsy) % nsy;
                    sb = wy * nsx + wx;
                    zx = (Lx * (qx - wx)) / nsx;
                    zy = (Ly * (qy - wy)) / nsy;
                }
                for (int64_t sk = os[sb]; sk < os[sb + 1]; ++sk) {
                    int64_t s = ps[sk];
                    double xs = psrc[2 * s], ys = psrc[2 * s + 1];
                    /* d = x_s - x_t + z with the forward sums' rounding */
                    double d[2];
                    if (dlp) {
                        d[0] = xs - tx + zx;
                        d[1] = ys - ty + zy;
                    } else if (own) {
                        d[0] = -(tx - xs);   /* forward: tx - xs; negation is exact */
                        d[1] = -(ty - ys);
                    } else {
                        d[0] = xs - (tx - zx);
                        d[1] = ys - (ty - zy);
                    }
                    double r2 = d[0] * d[0] + d[1] * d[1];
                    double *ms = mine + (size_t)W * s;
                    if (own && r2 == 0.0) {
                        if (!dlp) {   /* self term: constant in the positions */
                            ms[2] += self * a[0];
                            ms[3] += self * a[1];
                        }
                        continue;
                    }
                    if (!own && !(r2 < cut2)) continue;
                    double g[2] = {0.0, 0.0};
                    if (!dlp) {
                        double bvec[2] = {f[2 * s], f[2 * s + 1]};
                        double v[2] = {0.0, 0.0};
                        or_slp_pair_adj(d, r2, xi, a, bvec, v, g);
                        ms[2] += v[0];
                        ms[3] += v[1];
                    } else {
                        double S[2][2] = {{f[2 * s] * nv[2 * s], f[2 * s] * nv[2 * s + 1]},
                                          {f[2 * s + 1] * nv[2 * s], f[2 * s + 1] * nv[2 * s + 1]}};
                        double M[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
                        or_dlp_pair_adj(d, r2, xi, a, S, M, g);
                        ms[2] += M[0][0];
                        ms[3] += M[0][1];
                        ms[4] += M[1][0];
                        ms[5] += M[1][1];
                    }
                    ms[0] += g[0];
                    ms[1] += g[1];
                    gt[0] -= g[0];
                    gt[1] -= g[1];
                }
            }
            tacc[2 * t] = gt[0];
            tacc[2 * t + 1] = gt[1];
        }
    }
    const double s4 = 1.0 / (4 * OR_PI);
    if (dptar)
        for (int64_t t = 0; t < nt; ++t) {
            dptar[2 * t] = tacc[2 * t] * s4;
            dptar[2 * t + 1] = tacc[2 * t + 1] * s4;
        }
    for (int64_t s = 0; s < ns; ++s) {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int th = 0; th < nthr; ++th)
            for (int c = 0; c < W; ++c) acc[c] += sbuf[(size_t)th * (size_t)(W * ns + 1) + (size_t)W * s + c];
        for (int c = 0; c < W; ++c) acc[c] *= s4;
        if (dpsrc) {
            dpsrc[2 * s] = acc[0];
            dpsrc[2 * s + 1] = acc[1];
        }
        if (!dlp) {
            if (df) {
                df[2 * s] = acc[2];
                df[2 * s + 1] = acc[3];
            }
        } else {
            /* M_jk = dL/dS_jk with S_jk = f_j n_k */
            double n1 = nv[2 * s], n2 = nv[2 * s + 1], f1 = f[2 * s], f2 = f[2 * s + 1];
            if (df) {
                df[2 * s] = acc[2] * n1 + acc[3] * n2;
                df[2 * s + 1] = acc[4] * n1 + acc[5] * n2;
            }
            if (dn) {
                dn[2 * s] = acc[2] * f1 + acc[4] * f2;
                dn[2 * s + 1] = acc[3] * f1 + acc[5] * f2;
            }
        }
    }
    free(ps); free(pt); free(os); free(ot); free(sbuf); free(tacc);
    return 0;
}

int or_slp_real_adj(const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                    const double *f, double xi, int nsx, int nsy, double Lx, double Ly,
                    const double *du, double *dpsrc, double *dptar, double *df)
{
    return or_real_adj(0, psrc, ns, ptar, nt, f, NULL, xi, nsx, nsy, Lx, Ly, du, dpsrc,
                       dptar, df, NULL);
}

int or_dlp_real_adj(const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                    const double *f, const double *nv, double xi, int nsx, int nsy,
                    double Lx, double Ly, const double *du, double *dpsrc, double *dptar,
                    double *df, double *dn)
{
    return or_real_adj(1, psrc, ns, ptar, nt, f, nv, xi, nsx, nsy, Lx, Ly, du, dpsrc, dptar,
                       df, dn);
}

/* Interpolation of ng grids at n points with the forward's weights and
 * scaling, plus the derivative with respect to the point:
 *   val[k*ng + c]  = s * sum_nodes W G_c                 (if val)
 *   grad[2k + l]  += s * sum_c wt[k*ng + c] * dW/dx_l G_c (if grad)
 * W = exp(c |node - x|^2) is evaluated as the forward does (FGG factors);
 * dW/dx = -2c (h i - w - px) W for node index i of the (P+1)^2 support,
 * the derivative of the factored exponent c (px^2 + 2 w px - 2 h i px + ...). */
static void or_interp_vg(const double *pts, int64_t n, int ng, double *const *grids,
                         double xi, double eta, int Mx, int My, double Lx, double Ly, double w,
                         int P, const double *e1, double *val, const double *wt, double *grad)
{
    double h = Lx / Mx;
    double s = 4 * xi * xi / eta;
    s = s * s * h * h / OR_PI;
    s = s / (4 * OR_PI);
    double c = -2 * xi * xi / eta;
#pragma omp parallel for
    for (int64_t k = 0; k < n; ++k) {
        or_fgg_point g;
        double px, py;
        int mx, my;
        or_fgg_setup(pts[2 * k], pts[2 * k + 1], Lx, Ly, h, P, xi, eta, w, &g);
        or_find_closest_node(pts[2 * k], pts[2 * k + 1], Lx, Ly, h, P, &mx, &my, &px, &py);
        double v[8] = {0}, gx[8] = {0}, gy[8] = {0};
        double ex = g.ex;
        for (int i = 0; i <= P; ++i) {
            int col = ((i + g.mx) % Mx + Mx) % Mx;
            double ey = ex * g.e4y * e1[i];
            double dx = h * i - w - px;
            for (int j = 0; j <= P; ++j) {
                int row = ((j + g.my) % My + My) % My;
                double wgt = ey * e1[j];
                double dy = h * j - w - py;
                size_t idx = (size_t)col * My + row;
                for (int q = 0; q < ng; ++q) {
                    double gv = grids[q][idx];
                    v[q] += wgt * gv;
                    gx[q] += -2 * c * dx * wgt * gv;
                    gy[q] += -2 * c * dy * wgt * gv;
                }
                ey *= g.e3y;
            }
            ex *= g.e3x;
        }
        if (val)
            for (int q = 0; q < ng; ++q) val[k * ng + q] = s * v[q];
        if (grad) {
            double a0 = 0.0, a1 = 0.0;
            for (int q = 0; q < ng; ++q) {
                a0 += wt[k * ng + q] * gx[q];
                a1 += wt[k * ng + q] * gy[q];
            }
            grad[2 * k] = s * a0;
            grad[2 * k + 1] = s * a1;
        }
    }
}

/* Stokeslet Fourier-space adjoint.  The forward is u = s I_t Re(F^-1 B F S_s f)
 * with B the real symmetric 2x2 filter per wavenumber (F unnormalised DFT,
 * S spreading, I interpolation).  For a real-linear map X -> Re(Y X) the
 * adjoint is Z -> Re(Y^H Z), and Y^H = F^-1 B^H F = F^-1 B F, so
 * df = s S_s^T Re(F^-1 B F I_t^T du): the forward pipeline with the roles of
 * sources and targets exchanged and du as the density. */
int or_slp_kspace_adj(const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                      double xi, double eta, const double *f, int Mx, int My, double Lx,
                      double Ly, double w, int P, const double *du, double *dpsrc,
                      double *dptar, double *df)
{
    size_t G = (size_t)Mx * My;
    double h = Lx / Mx;
    double *e1 = or_e1_table(xi, eta, h, P);
    double *A[2], *B[2];
    for (int c = 0; c < 2; ++c) {
        A[c] = (double *)calloc(G, sizeof(double));
        B[c] = (double *)calloc(G, sizeof(double));
    }
    if (dptar) {
        or_slp_kspace_grids(psrc, ns, xi, eta, f, Mx, My, Lx, Ly, w, P, e1, A[0], A[1]);
        or_interp_vg(ptar, nt, 2, A, xi, eta, Mx, My, Lx, Ly, w, P, e1, NULL, du, dptar);
    }
    if (dpsrc || df) {
        or_slp_kspace_grids(ptar, nt, xi, eta, du, Mx, My, Lx, Ly, w, P, e1, B[0], B[1]);
        double *val = (double *)malloc(sizeof(double) * (size_t)(2 * ns + 2));
        or_interp_vg(psrc, ns, 2, B, xi, eta, Mx, My, Lx, Ly, w, P, e1, val, f, dpsrc);
        if (df) memcpy(df, val, sizeof(double) * (size_t)(2 * ns));
        free(val);
    }
    for (int c = 0; c < 2; ++c) { free(A[c]); free(B[c]); }
    free(e1);
    return 0;
}

/* Stresslet Fourier-space adjoint.  The forward is u = s I_t Im(F^-1 B F S_s q)
 * with q the four densities n_l f_j and B the real 2x4 filter.  For
 * X -> Im(Y X) = Re(-i Y X) the adjoint is Z -> Re(i Y^H Z) = -Im(Y^H Z), with
 * Y^H = F^-1 B^T F: spread du at the targets, apply the transposed filter
 * (2 -> 4 grids), inverse transform, keep -Im, interpolate at the sources. */
int or_dlp_kspace_adj(const double *psrc, int64_t ns, const double *ptar, int64_t nt,
                      double xi, double eta, const double *f, const double *nv, int Mx,
                      int My, double Lx, double Ly, double w, int P, const double *du,
                      double *dpsrc, double *dptar, double *df, double *dn)
{
    size_t G = (size_t)Mx * My;
    double h = Lx / Mx;
    double *e1 = or_e1_table(xi, eta, h, P);
    if (dptar) {
        double *A[2];
        for (int c = 0; c < 2; ++c) A[c] = (double *)calloc(G, sizeof(double));
        or_dlp_kspace_grids(psrc, ns, xi, eta, f, nv, Mx, My, Lx, Ly, w, P, e1, A[0], A[1]);
        or_interp_vg(ptar, nt, 2, A, xi, eta, Mx, My, Lx, Ly, w, P, e1, NULL, du, dptar);
        for (int c = 0; c < 2; ++c) free(A[c]);
    }
    if (dpsrc || df || dn) {
        double *Zr[2], *Zi[2], *Or[4], *Oi[4];
        for (int c = 0; c < 2; ++c) {
            Zr[c] = (double *)calloc(G, sizeof(double));
            Zi[c] = (double *)calloc(G, sizeof(double));
        }
        for (int c = 0; c < 4; ++c) {
            Or[c] = (double *)calloc(G, sizeof(double));
            Oi[c] = (double *)calloc(G, sizeof(double));
        }
        or_spread(ptar, nt, du, 2, xi, eta, Mx, My, Lx, Ly, w, P, e1, Zr);
        or_fft2(Zr[0], Zi[0], Mx, My, -1);
        or_fft2(Zr[1], Zi[1], Mx, My, -1);
#pragma omp parallel for
        for (int j = 0; j < Mx; ++j) {
            double k1 = or_wavenumber(j, Mx, Lx);
            for (int k = 0; k < My; ++k) {
                double k2 = or_wavenumber(k, My, Ly);
                double et, b111, b112, b122, b222;
                or_dlp_filter_coef(k1, k2, xi, eta, &et, &b111, &b112, &b122, &b222);
                size_t p = (size_t)j * My + k;
                /* B^T: column c of the forward's 2x4 matrix
                 * [b111 b112 b112 b122; b112 b122 b122 b222] */
                double col[4][2] = {{b111, b112}, {b112, b122}, {b112, b122}, {b122, b222}};
                for (int c = 0; c < 4; ++c) {
                    Or[c][p] = et * (col[c][0] * Zr[0][p] + col[c][1] * Zr[1][p]);
                    Oi[c][p] = et * (col[c][0] * Zi[0][p] + col[c][1] * Zi[1][p]);
                }
            }
        }
        for (int c = 0; c < 4; ++c) {
            Or[c][0] = Oi[c][0] = 0.0;   /* k = 0 mode, as in the forward */
            or_fft2(Or[c], Oi[c], Mx, My, +1);
            for (size_t i = 0; i < G; ++i) Oi[c][i] = -Oi[c][i];
        }
        double *q = or_dlp_q4(f, nv, ns);
        double *val = (double *)malloc(sizeof(double) * (size_t)(4 * ns + 4));
        or_interp_vg(psrc, ns, 4, Oi, xi, eta, Mx, My, Lx, Ly, w, P, e1, val, q, dpsrc);
        /* val = dL/d(n1 f1, n2 f1, n1 f2, n2 f2) */
        for (int64_t s = 0; s < ns; ++s) {
            const double *D = val + 4 * s;
            double n1 = nv[2 * s], n2 = nv[2 * s + 1], f1 = f[2 * s], f2 = f[2 * s + 1];
            if (df) {
                df[2 * s] = D[0] * n1 + D[1] * n2;
                df[2 * s + 1] = D[2] * n1 + D[3] * n2;
            }
            if (dn) {
                dn[2 * s] = D[0] * f1 + D[2] * f2;
                dn[2 * s + 1] = D[1] * f1 + D[3] * f2;
            }
        }
        free(q);
        free(val);
        for (int c = 0; c < 2; ++c) { free(Zr[c]); free(Zi[c]); }
        for (int c = 0; c < 4; ++c) { free(Or[c]); free(Oi[c]); }
    }
    free(e1);
    return 0;
}

/* Thread control for the CPU baseline. */
void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
