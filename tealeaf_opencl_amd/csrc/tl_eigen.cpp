// Host scalar routines of the Chebyshev / PPCG set-up (tea_leaf_cheby.f90:
// 44-172): implicit QL eigenvalues of the Lanczos tridiagonal built from the
// CG coefficients, and the Chebyshev recurrence coefficients.
#include "tl_eigen.h"

#include <math.h>

#include <algorithm>

namespace tl {


// implicit QL on a symmetric tridiagonal matrix (tea_leaf_cheby.f90:44-94)
int tqli(std::vector<double> &d, std::vector<double> &e, int n)
{
    for (int i = 0; i < n - 1; i++) e[i] = e[i + 1];
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m < n - 1; m++) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (iter == 30) return 1;
            iter++;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = sqrt(g * g + 1.0 * 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool restart = false;
            for (int i = m - 1; i >= l; i--) {
                const double f = s * e[i];
                const double b = c * e[i];
                r = sqrt(f * f + g * g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] = d[i + 1] - p;
                    e[m] = 0.0;
                    restart = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
            }
            if (restart) continue;
            d[l] = d[l] - p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

// tea_calc_eigenvalues (tea_leaf_cheby.f90:96-143); alphas/betas 0-based
int calc_eigenvalues(const double *a, const double *b, int ns, double *eigmin, double *eigmax)
{
    if (ns < 1) return 1;
    std::vector<double> d(ns, 0.0), e(ns, 0.0);
    for (int n = 0; n < ns; n++) {
        d[n] = 1.0 / a[n];
        if (n > 0) d[n] = d[n] + b[n - 1] / a[n - 1];
        if (n < ns - 1) e[n + 1] = sqrt(b[n]) / a[n];
    }
    int info = tqli(d, e, ns);
    if (info) return info;
    bool swapped;
    do {
        swapped = false;
        for (int n = 0; n < ns - 1; n++)
            if (d[n] > d[n + 1]) {
                std::swap(d[n], d[n + 1]);
                swapped = true;
            }
    } while (swapped);
    *eigmin = d[0];
    *eigmax = d[ns - 1];
    if (*eigmin < 0.0 || *eigmax < 0.0) info = 1;
    return info;
}

// tea_calc_ch_coefs (tea_leaf_cheby.f90:145-172)
void calc_ch_coefs(std::vector<double> &ca, std::vector<double> &cb, double eigmin, double eigmax, double *theta,
                   int n)
{
    const double th = (eigmax + eigmin) / 2;
    const double delta = (eigmax - eigmin) / 2;
    const double sigma = th / delta;
    double rho_old = 1.0 / sigma;
    ca.assign(n > 0 ? n : 1, 0.0);
    cb.assign(n > 0 ? n : 1, 0.0);
    for (int i = 0; i < n; i++) {
        const double rho_new = 1.0 / (2.0 * sigma - rho_old);
        ca[i] = rho_new * rho_old;
        cb[i] = 2.0 * rho_new / delta;
        rho_old = rho_new;
    }
    *theta = th;
}

}  // namespace tl
