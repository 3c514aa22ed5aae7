"""The library's host routines between the CG presteps and the Chebyshev /
PPCG phase (tealeaf_opencl_amd/csrc/tl_eigen.cpp: tqli, tea_calc_eigenvalues
and tea_calc_ch_coefs of tea_leaf_cheby.f90:44-172), on the CPU: bitwise
against the oracle's restatement (oracle/tl_oracle.c, written apart), and
against what mathematics fixes: LAPACK's eigenvalues of the same Lanczos
matrix, the spectrum a CG run on a known diagonal operator recovers, and the
closed form of the Chebyshev recurrence."""
import math

import numpy as np
import pytest


def cg_coefficients(lam, steps):
    """alphas, betas of plain CG on diag(lam) from b = 1."""
    r = np.ones_like(lam)
    p = r.copy()
    rro = r @ r
    alphas, betas = [], []
    for _ in range(steps):
        w = lam * p
        a = rro / (p @ w)
        r = r - a * w
        rrn = r @ r
        alphas.append(a)
        betas.append(rrn / rro)
        p = r + (rrn / rro) * p
        rro = rrn
    return np.array(alphas), np.array(betas)


def lanczos_matrix(a, b):
    """tea_calc_eigenvalues' tridiagonal from CG's alphas and betas."""
    n = len(a)
    d = np.array([1.0 / a[i] + (b[i - 1] / a[i - 1] if i > 0 else 0.0) for i in range(n)])
    e = np.array([math.sqrt(b[i]) / a[i] for i in range(n - 1)])
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


@pytest.mark.parametrize("steps", [2, 5, 20, 60])
def test_eigenvalue_bounds(hiplib, oracle, steps):
    lam = np.linspace(1.0, 100.0, 200)
    a, b = cg_coefficients(lam, steps)
    lo, hi = hiplib.calc_eigenvalues(a, b)
    info, olo, ohi = oracle.calc_eigenvalues(a, b)
    assert info == 0 and (lo, hi) == (olo, ohi)
    ev = np.linalg.eigvalsh(lanczos_matrix(a, b))
    assert abs(lo - ev[0]) <= 1e-10 * ev[-1] and abs(hi - ev[-1]) <= 1e-10 * ev[-1]
    if steps == 60:
        assert abs(lo - 1.0) < 1e-6 and abs(hi - 100.0) < 1e-6


def test_eigenvalue_bounds_seeded_spectra(hiplib, oracle):
    rng = np.random.default_rng(7)
    for _ in range(20):
        lam = np.sort(10.0 ** rng.uniform(-1, 3, 300))
        a, b = cg_coefficients(lam, int(rng.integers(3, 40)))
        got = hiplib.calc_eigenvalues(a, b)
        info, olo, ohi = oracle.calc_eigenvalues(a, b)
        assert info == 0 and got == (olo, ohi)


@pytest.mark.parametrize("eigmin,eigmax,n", [(0.3, 47.0, 40), (1.0, 1.0e4, 10), (0.95, 1.05, 7)])
def test_chebyshev_coefficients(hiplib, oracle, eigmin, eigmax, n):
    a, b, theta = hiplib.calc_ch_coefs(eigmin, eigmax, n)
    oa, ob, otheta = oracle.calc_ch_coefs(eigmin, eigmax, n)
    assert np.array_equal(a, oa) and np.array_equal(b, ob) and theta == otheta
    th, delta = 0.5 * (eigmax + eigmin), 0.5 * (eigmax - eigmin)
    sigma = th / delta
    T = lambda m: math.cosh(m * math.acosh(sigma))
    rho = [T(m) / T(m + 1) for m in range(n + 1)]
    assert theta == th
    for i in range(n):
        assert abs(a[i] - rho[i + 1] * rho[i]) <= 1e-12 * abs(a[i])
        assert abs(b[i] - 2.0 * rho[i + 1] / delta) <= 1e-12 * abs(b[i])


def test_bad_arguments(hiplib):
    with pytest.raises(hiplib.TealeafError):
        hiplib.calc_eigenvalues(np.array([]), np.array([]))
    with pytest.raises(hiplib.TealeafError):
        hiplib.calc_ch_coefs(1.0, 2.0, 0)
