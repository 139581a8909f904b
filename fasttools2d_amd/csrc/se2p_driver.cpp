// se2p_driver.cpp -- the reference drivers as single C entry points:
// parameter selection (StokesSLP_ewald_2p.m:27-104, StokesDLP_ewald_2p.m:28-125),
// wrapping into [-L/2, L/2) (:75-78 / :94-97), the real- and Fourier-space
// sums, and the composition (:133-136; DLP zero mode and sign :150-158).
#include <cmath>
#include <cstring>

#include "se2p_device.h"
#include "se2p_internal.h"

namespace se2p {

static double mround(double x) { return std::copysign(std::floor(std::fabs(x) + 0.5), x); }

// MATLAB rat: continued fraction with round-to-nearest steps, tol 1e-6*|x|.
static void mat_rat(double x, long long *num, long long *den)
{
    const double tol = 1e-6 * std::fabs(x);
    double lastn = 1, lastd = 0, n = mround(x), d = 1, frac = x - n;
    int guard = 0;
    while (std::fabs(x - n / d) >= tol && guard++ < 64) {
        const double flip = 1.0 / frac;
        const double step = mround(flip);
        frac = flip - step;
        const double sn = n, sd = d;
        n = n * step + lastn;
        d = d * step + lastd;
        lastn = sn;
        lastd = sd;
    }
    if (d < 0) {
        n = -n;
        d = -d;
    }
    *num = (long long)n;
    *den = (long long)d;
}

template <class F, class DF>
static bool newton(F f, DF df, double x0, double tol, double *x_out)
{
    double x = x0;
    int it = 0;
    bool ok = true;
    while (std::fabs(f(x)) > tol) {
        if (it > 100) {
            ok = false;   // the reference prints a warning and keeps the iterate
            break;
        }
        const double step = std::min(f(x) / df(x), 0.9 * x);
        x -= step;
        ++it;
    }
    *x_out = x;
    return ok;
}

// Fixed-grid mode (p->fixed_grid, p->M > 0): the drivers' k-space estimate
// inverted for xi at kinf = M/2, the real-space estimate inverted for the
// cutoff at that xi, cells at least that wide.  Closed forms except the
// stresslet's real-space estimate, x e^{-x^2} s = tol with x = xi rc, solved
// by Newton on x^2 - ln x = ln(s / tol) on its decreasing branch.
static int select_params_grid(int kind, int64_t ns, int64_t nt, double Q, double Lx, double Ly,
                              se2p_params *p)
{
    const int P = p->P > 0 ? p->P : 24;
    const double tol = p->tol > 0 ? p->tol : 1e-16;
    const int Mb = p->M;
    long long A, B;
    mat_rat(Lx / Ly, &A, &B);
    const double kinf = Mb / 2;
    const double k = 2 * SE2P_PI * kinf / Lx;
    double xi, rc_min;
    bool ok = true;
    if (kind == 0) {
        const double c = 4 * Q * SE2P_PI * Lx / (Lx * Lx * Lx * Lx * Lx * Lx);
        xi = k / (2 * std::sqrt(std::log(std::sqrt(c / k) / tol)));
        rc_min = std::sqrt(std::log(std::sqrt(Q * SE2P_PI / (4 * Lx * Ly * xi)) / tol)) / xi;
    } else {
        const double c = 8 * SE2P_PI * Q * Lx / (Lx * Lx * Lx) * (Lx * Lx * Lx);
        xi = k / (2 * std::sqrt(std::log(std::sqrt(k * c) / tol)));
        const double b = std::log(std::sqrt(2 * SE2P_PI * Q / (Lx * Ly)) / tol);
        double x = std::sqrt(std::max(b, 0.5)) + 1.0;
        ok = false;
        for (int it = 0; it < 100; ++it) {
            const double g = x * x - std::log(x) - b;
            const double xn = x - g / (2 * x - 1 / x);
            if (std::fabs(xn - x) <= 1e-15 * x) {
                x = xn;
                ok = true;
                break;
            }
            x = std::max(xn, 0.75);
        }
        rc_min = x / xi;
    }
    if (!(xi > 0) || !std::isfinite(xi) || !(rc_min > 0) || !std::isfinite(rc_min))
        return SE2P_ERR_PARAM;   // tol too loose for this grid (estimate below tol everywhere)
    const long long a = std::max(1LL, (long long)std::floor(Lx / (A * rc_min) * (1 + 1e-12)));
    const double m = 0.95 * std::sqrt(SE2P_PI * P);
    p->P = P;
    p->tol = tol;
    p->nside_x = (int)(a * A);
    p->nside_y = (int)(a * B);
    p->Nb = (double)(ns + nt) / ((double)p->nside_x * p->nside_y);
    p->rc = Lx / p->nside_x;
    p->xi = xi;
    p->kinf = (int)kinf;
    p->Mx = (int)(A * Mb);
    p->My = (int)(B * Mb);
    p->w = P * Lx / p->Mx / 2;
    p->eta = std::pow(2 * xi * p->w / m, 2);
    p->converged = ok ? 1 : 0;
    return 0;
}

int select_params(int kind, int64_t ns, int64_t nt, double Q, double Lx, double Ly,
                  se2p_params *p)
{
    if (p->fixed_grid) return select_params_grid(kind, ns, nt, Q, Lx, Ly, p);
    const int P = p->P > 0 ? p->P : 24;
    const double npts = (double)(ns + nt);
    double Nb = p->Nb;
    if (!(Nb > 0)) Nb = std::min(P * mround(std::log2(npts)), npts / 4);
    const double tol = p->tol > 0 ? p->tol : 1e-16;
    long long A, B;
    mat_rat(Lx / Ly, &A, &B);
    const long long a = (long long)std::ceil(std::sqrt(npts / (Nb * A * B)));
    const double m = 0.95 * std::sqrt(SE2P_PI * P);
    const int nsx = (int)(a * A), nsy = (int)(a * B);
    const double rc = Lx / nsx;
    double xi, k;
    bool ok1, ok2;
    if (kind == 0) {   // StokesSLP_ewald_2p.m:147-203
        const double s = Q * SE2P_PI / (4 * Lx * Ly);
        ok1 = newton([&](double x) { return std::sqrt(s / x) * std::exp(-x * x * rc * rc) - tol; },
                     [&](double x) {
                         return -std::sqrt(s) * std::exp(-x * x * rc * rc) *
                                (0.5 * std::pow(x, -1.5) + 2 * rc * rc * x * std::sqrt(1 / x));
                     },
                     4 / rc, tol, &xi);
        const double Lk = Lx;   // the driver passes Lx twice
        const double c = 4 * Q * SE2P_PI * std::max(Lx, Lk) / (Lx * Lx * Lx * Lk * Lk * Lk);
        ok2 = newton([&](double kk) { return std::sqrt(c / kk) * std::exp(-kk * kk / (4 * xi * xi)) - tol; },
                     [&](double kk) {
                         return -std::sqrt(c) * std::exp(-kk * kk / (4 * xi * xi)) *
                                (0.5 * std::pow(kk, -1.5) + std::sqrt(1 / kk) * 2 * kk / (4 * xi * xi));
                     },
                     mround(5 * xi), tol, &k);
    } else {           // StokesDLP_ewald_2p.m:171-225
        const double s = std::sqrt(2 * SE2P_PI * Q / (Lx * Ly));
        ok1 = newton([&](double x) { return std::exp(-x * x * rc * rc) * x * rc * s - tol; },
                     [&](double x) { return (1 - 2 * x * x * rc * rc) * std::exp(-x * x * rc * rc) * s * rc; },
                     1 / rc, tol, &xi);
        const double Lk = Lx;
        // MATLAB precedence: .../Lx^3*Ly^3 (division binds to Lx^3 only)
        const double c = 8 * SE2P_PI * Q * std::max(Lx, Lk) / (Lx * Lx * Lx) * (Lk * Lk * Lk);
        ok2 = newton([&](double kk) { return std::sqrt(kk * c) * std::exp(-kk * kk / (4 * xi * xi)) - tol; },
                     [&](double kk) {
                         return std::sqrt(c) * std::exp(-kk * kk / (4 * xi * xi)) *
                                (0.5 * std::pow(kk, -0.5) - 2 * std::pow(kk, 1.5) / (4 * xi * xi));
                     },
                     mround(5 * xi), tol, &k);
    }
    const long long kinf = (long long)mround(Lx * k / (2 * SE2P_PI));
    const long long Mb = p->M > 0 ? p->M : std::min(2 * kinf, 10000LL);
    p->P = P;
    p->Nb = Nb;
    p->tol = tol;
    p->nside_x = nsx;
    p->nside_y = nsy;
    p->rc = rc;
    p->xi = xi;
    p->kinf = (int)kinf;
    p->My = (int)(B * Mb);
    p->Mx = (int)(A * Mb);
    p->w = P * Lx / p->Mx / 2;
    p->eta = std::pow(2 * xi * p->w / m, 2);
    p->converged = (ok1 && ok2) ? 1 : 0;
    return 0;
}

}  // namespace se2p
