"""Helpers that set up the same chunk on the GPU library and on the CPU oracle."""
import numpy as np

from tealeaf_opencl_amd.decks import seeded_states

GEOM = {"rectangle": 1, "circle": 2, "point": 3}


def state_arrays(states, x_cells, y_cells, xmax=10.0, ymax=10.0):
    """Deck states -> the arrays generate_chunk takes, with the read_input.f90:320-327 shrink."""
    n = len(states)
    dens = np.array([s["density"] for s in states], dtype=np.float64)
    ener = np.array([s["energy"] for s in states], dtype=np.float64)
    xmin = np.array([s.get("xmin", 0.0) for s in states], dtype=np.float64)
    xmx = np.array([s.get("xmax", 0.0) for s in states], dtype=np.float64)
    ymin = np.array([s.get("ymin", 0.0) for s in states], dtype=np.float64)
    ymx = np.array([s.get("ymax", 0.0) for s in states], dtype=np.float64)
    rad = np.array([s.get("radius", 0.0) for s in states], dtype=np.float64)
    geom = np.array([GEOM.get(s.get("geometry", ""), 0) for s in states], dtype=np.int32)
    dx = xmax / float(np.float32(x_cells))
    dy = ymax / float(np.float32(y_cells))
    for i in range(1, n):
        xmin[i] += dx / 100.0
        ymin[i] += dy / 100.0
        xmx[i] -= dx / 100.0
        ymx[i] -= dy / 100.0
    return dens, ener, xmin, xmx, ymin, ymx, rad, geom


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
