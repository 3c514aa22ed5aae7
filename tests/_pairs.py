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
