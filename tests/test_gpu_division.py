"""The division the iteration kernels use for recomputed face coefficients,
block-Jacobi factors and diagonals (div_nr: hardware reciprocal, two Newton
steps, one residual correction) against the IEEE FP64 quotient, bit for bit,
over the operand kinds those kernels meet: face numerators and denominators
w_a + w_b, 2 w_a w_b of conductivities 1e-3..1e4; reciprocals of Thomas
pivots 1..2e6; residuals of either sign 1e-40..1e10 and negative faces over
pivots; zero dividends. The GPU's IEEE quotient is also checked against
numpy's (the host's) division. tl_common_init refuses densities outside
[2^-499, 2^499]; the divisors densities inside it give (face denominators
down to 2^-997 and up to 2^999, pivots up to 2^540) are checked at the edges
of that range too."""
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


def test_division_bitwise_ieee_at_the_range_edges(hiplib):
    """Face numerators and denominators of conductivities at and near 2^-499
    and 2^499 (the range tl_common_init accepts), and pivots 1 + faces with
    faces up to rx * 2^499 for rx up to 2^40."""
    rng = np.random.default_rng(7)
    k = 1 << 18
    edge = np.array([2.0 ** -499, 2.0 ** 499, 2.0 ** -498.5, 2.0 ** 498.5])
    wa = np.concatenate([np.repeat(edge, 4), 2.0 ** rng.uniform(-499, 499, k)])
    wb = np.concatenate([np.tile(edge, 4), 2.0 ** rng.uniform(-499, 499, k)])
    faces = 2.0 ** rng.uniform(0, 40, wa.size) * (wa + wb) / (2.0 * wa * wb)
    num = np.concatenate([wa + wb, -faces, np.ones(wa.size)])
    den = np.concatenate([2.0 * wa * wb, 1.0 + 4.0 * faces, 1.0 + 2.0 * faces])
    assert np.all(np.isfinite(den)) and np.all(den > 0)
    quot, fast = hiplib.check_division(num, den)
    assert np.array_equal(quot.view(np.uint64), (num / den).view(np.uint64))
    bad = np.nonzero(quot.view(np.uint64) != fast.view(np.uint64))[0]
    assert bad.size == 0, [(num[i], den[i], quot[i], fast[i]) for i in bad[:5]]


@pytest.mark.parametrize("value,accepted", [(0.0, False), (1e-160, False), (1e160, False), (float("nan"), False),
                                            (-1.0, False), (1e-140, True), (1e140, True)])
def test_common_init_density_range(hiplib, value, accepted):
    """An interior density outside [2^-499, 2^499] (or not a number) makes
    tl_common_init fail with a message; inside, it succeeds. The zero halos
    of a fresh chunk are not interior and do not count."""
    from tealeaf_opencl_amd import decks
    n = 64
    c = hiplib.Chunk(n, n, 1, hiplib.PREC_JAC_BLOCK)
    try:
        c.initialise_chunk(0.0, 0.0, 10.0 / n, 10.0 / n)
        c.generate_chunk(*decks.state_arrays(decks.seeded_states(1, 2, False), n, n))
        c.set_field()
        c.common_init(1, 0.004)  # the painted densities pass
        rho = c.get_field("density")
        rho[20, 30] = value
        c.put_field("density", rho)
        for coef in (1, 2):
            if accepted:
                c.common_init(coef, 0.004)
            else:
                with pytest.raises(hiplib.TealeafError, match="outside"):
                    c.common_init(coef, 0.004)
    finally:
        c.close()
