"""tl_generate_chunk_noalloc (no reference counterpart: generate_chunk for
callers that cannot allocate device memory) against tl_generate_chunk, the
entry point whose per-call device staging it replaces with a file-scope
device array: bitwise the same density, energy0, u and u0, on two calls in a
row through the C-ABI (the second with fewer states, on a smaller chunk), and
a clear error for more states than the array holds."""
import numpy as np
import pytest

from tealeaf_opencl_amd import decks

pytestmark = pytest.mark.gpu

FIELDS = ("density", "energy0", "u", "u0")


def _chunk(hiplib, x, y, hd=1):
    c = hiplib.Chunk(x, y, hd, 1)
    c.initialise_chunk(0.0, 0.0, 10.0 / x, 10.0 / y)
    return c


@pytest.mark.parametrize("order", ["noalloc_first", "reference_first"])
def test_noalloc_matches_generate_chunk(hiplib, order):
    cases = [(200, 148, 1, decks.seeded_states(3, 6, True)), (67, 13, 2, decks.seeded_states(5, 1, False))]
    assert len(cases[1][3]) < len(cases[0][3])
    for x, y, hd, states in cases:
        arrays = decks.state_arrays(states, x, y)
        got, ref = _chunk(hiplib, x, y, hd), _chunk(hiplib, x, y, hd)
        runs = [(got, True), (ref, False)]
        if order == "reference_first":
            runs.reverse()
        for c, noalloc in runs:
            c.generate_chunk(*arrays, noalloc=noalloc)
        for f in FIELDS:
            a, b = got.get_field(f), ref.get_field(f)
            assert np.array_equal(a, b), (x, y, f)
        # the painted states are really there (not a field left at zero)
        dens = got.interior(got.get_field("density"))
        assert set(np.unique(dens)) <= {s["density"] for s in states} and np.all(dens > 0)
        got.close()
        ref.close()


def test_noalloc_rejects_too_many_states(hiplib):
    c = _chunk(hiplib, 64, 64)
    n = hiplib.NOALLOC_MAX_STATES + 1
    arrays = [np.ones(n)] * 7 + [np.ones(n, dtype=np.int32)]
    with pytest.raises(hiplib.TealeafError, match="states outside"):
        c.generate_chunk(*arrays, noalloc=True)
    # the chunk is still usable afterwards
    c.generate_chunk(*decks.state_arrays(decks.seeded_states(1, 2, False), 64, 64), noalloc=True)
    assert np.all(c.interior(c.get_field("density")) > 0)
    c.close()
