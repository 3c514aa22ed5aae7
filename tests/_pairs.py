"""Helpers that set up the same chunk on the GPU library and on the CPU oracle."""
import numpy as np

from tealeaf_opencl_amd.decks import seeded_states, state_arrays

def setup(chunk, seed=1, n_rect=2, circle=True, dt=0.004, coef=1, reflective=False, zero=(1, 1, 1, 1)):
    """initialise_chunk + generate_chunk + set_field + common_init on one chunk."""
    x, y = chunk.x_max, chunk.y_max
    dx, dy = 10.0 / x, 10.0 / y
    chunk.initialise_chunk(0.0, 0.0, dx, dy)
    chunk.generate_chunk(*state_arrays(seeded_states(seed, n_rect, circle), x, y))
    if reflective:
        chunk.update_halo((1, 1, 1), chunk.hd)
    chunk.set_field()
    return chunk.common_init(coef, dt, zero_boundary=zero, reflective_boundary=reflective)


def rand_interior(chunk, rng, scale=1.0):
    """A random field that is zero in the halo (as every solver vector is)."""
    a = np.zeros(chunk.shape)
    chunk.interior(a)[:] = scale * rng.standard_normal((chunk.y_max, chunk.x_max))
    return a


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)


def scal_err(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def oracle_residual(o, u):
    """||u0 - A u||_2 over the interior, formed by the oracle."""
    o.put_field("u", u)
    o.calc_residual()
    return float(np.linalg.norm(o.interior(o.get_field("r"))))


def oracle_cg(o, tol=1e-14, max_iters=20000):
    """Plain FP64 PCG with the oracle's reference-granularity kernels from the
    oracle's current u to ||r|| <= tol ||r_0|| (tea_solve.f90:103-125, 270-300);
    returns u."""
    o.calc_residual()
    rro = o.cg_init()
    r0 = np.sqrt(o.calc_2norm(1))
    if r0 == 0.0:
        return o.get_field("u")
    for _ in range(max_iters):
        pw = o.cg_calc_w()
        alpha = rro / pw
        rrn = o.cg_calc_ur(alpha)
        if np.sqrt(o.calc_2norm(1)) <= tol * r0:
            break
        o.cg_calc_p(rrn / rro)
        rro = rrn
    return o.get_field("u")
