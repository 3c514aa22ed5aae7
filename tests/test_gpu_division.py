"""The division the iteration kernels use for recomputed face coefficients,
block-Jacobi factors and diagonals (div_nr: hardware reciprocal, two Newton
steps, one residual correction) against the IEEE FP64 quotient, bit for bit,
over the operand kinds those kernels meet: face numerators and denominators
w_a + w_b, 2 w_a w_b of conductivities 1e-3..1e4; reciprocals of Thomas
pivots 1..2e6; residuals of either sign 1e-40..1e10 and negative faces over
pivots; zero dividends. The GPU's IEEE quotient is also checked against
numpy's (the host's) division."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def operands(n, seed):
    rng = np.random.default_rng(seed)
    k = n // 4
    wa, wb = 10.0 ** rng.uniform(-3, 4, k), 10.0 ** rng.uniform(-3, 4, k)
    piv = lambda m: 10.0 ** rng.uniform(0, 6, m) * (1.0 + rng.uniform(0, 1, m))
    num = np.concatenate([wa + wb, np.ones(k),
                          rng.choice([-1.0, 1.0], k) * 10.0 ** rng.uniform(-40, 10, k),
                          -(10.0 ** rng.uniform(-4, 4, k)), [0.0, 0.0, 7.0, 1.0]])
    den = np.concatenate([2.0 * wa * wb, piv(k), piv(k), piv(k), [3.0, 1e6, 7.0, 1.0]])
    return num, den


@pytest.mark.parametrize("seed", [1, 2])
def test_division_bitwise_ieee(hiplib, seed):
    num, den = operands(1 << 22, seed)
    quot, fast = hiplib.check_division(num, den)
    assert np.array_equal(quot.view(np.uint64), (num / den).view(np.uint64))
    bad = np.nonzero(quot.view(np.uint64) != fast.view(np.uint64))[0]
    assert bad.size == 0, [(num[i], den[i], quot[i], fast[i]) for i in bad[:5]]
