// Host scalar routines of the Chebyshev / PPCG set-up (tea_leaf_cheby.f90:
// 44-172), shared by the native driver and the refined solve.
#pragma once

#include <vector>

namespace tl {

// implicit QL on a symmetric tridiagonal (diagonal d, off-diagonal e[1..n-1]);
// 0 on success, 1 after 30 sweeps without convergence
int tqli(std::vector<double> &d, std::vector<double> &e, int n);
// smallest and largest eigenvalue of the Lanczos tridiagonal of ns CG steps
// (alphas a[0..ns-1], betas b[0..ns-1]); nonzero on failure or a negative bound
int calc_eigenvalues(const double *a, const double *b, int ns, double *eigmin, double *eigmax);
// n Chebyshev coefficients (alphas ca, betas cb) for [eigmin, eigmax]; theta
// is the centre of the interval
void calc_ch_coefs(std::vector<double> &ca, std::vector<double> &cb, double eigmin, double eigmax, double *theta,
                   int n);

}  // namespace tl
