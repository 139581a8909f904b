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

int select_params(int kind, int64_t ns, int64_t nt, double Q, double Lx, double Ly,
                  se2p_params *p)
{
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
