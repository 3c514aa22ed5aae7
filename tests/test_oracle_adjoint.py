"""The density adjoint's derivation, pinned on the CPU: the oracle-based
reference gradient (tests/_adjoint_ref.py) against central finite
differences of the oracle's own forward timesteps, on a small grid with a
log-uniform density (both coefficient forms)."""
import numpy as np
import pytest

import _adjoint_ref as ref
from tealeaf_opencl_amd import decks


def problem(x, y, seed):
    rng = np.random.default_rng(seed)
    rho = np.zeros((y + 2, x + 2))
    rho[1:-1, 1:-1] = 10.0 ** rng.uniform(-1.0, 2.0, (y, x))
    e0 = np.zeros_like(rho)
    e0[1:-1, 1:-1] = rng.uniform(0.5, 2.0, (y, x))
    u_obs = np.zeros_like(rho)
    u_obs[1:-1, 1:-1] = rng.uniform(0.0, 50.0, (y, x))
    return rho, e0, u_obs


@pytest.mark.parametrize("coef", [1, 2])
@pytest.mark.parametrize("prec", [1, 3])
def test_adjoint_matches_finite_differences(oracle, coef, prec):
    x, y, nsteps, dt = 12, 10, 3, 0.004
    rho, e0, u_obs = problem(x, y, 5)
    o = ref.oracle_chunk(oracle, rho, e0, x, y, prec=prec)
    L, grad = ref.adjoint_grad(o, coef, dt, nsteps, u_obs)
    assert np.all(grad[0] == 0) and np.all(grad[:, 0] == 0)
    rng = np.random.default_rng(1)
    cells = [(1 + rng.integers(y), 1 + rng.integers(x)) for _ in range(12)] + [(1, 1), (y, x), (1, x)]
    for (k, j) in cells:
        h = 1e-6 * rho[k, j]
        vals = []
        for s in (1.0, -1.0):
            r2 = rho.copy()
            r2[k, j] += s * h
            o2 = ref.oracle_chunk(oracle, r2, e0, x, y, prec=prec)
            vals.append(ref.loss(o2, coef, dt, nsteps, u_obs))
        fd = (vals[0] - vals[1]) / (2.0 * h)
        assert abs(fd - grad[k, j]) <= 1e-6 * abs(grad[k, j]) + 1e-9 * L / rho[k, j], (k, j, fd, grad[k, j])


def test_config4_density_field():
    rho = decks.config4_density(64, seed=3)
    inner = rho[1:-1, 1:-1]
    assert rho.shape == (66, 66) and np.all(rho[0] == 0) and np.all(rho[:, -1] == 0)
    # 10^U[-1,2] times a bump in [1, 1.5]
    assert inner.min() >= 0.1 and inner.max() <= 150.0
    assert np.array_equal(rho, decks.config4_density(64, seed=3))
    assert not np.array_equal(rho, decks.config4_density(64, seed=4))
    r = decks.SplitMix64(99)
    assert np.array_equal(np.array([r.uniform() for _ in range(7)]), decks._splitmix_uniform(99, 7))
