"""Writes the FP64 CPU oracle's results for the 4096x4096 decks as fixtures:
the field summary after every timestep and the solve lines (iterations,
error). "cg", "chebyshev", "ppcg": the deck with two seeded rectangles and a
circle (seed 1, strip-4 block Jacobi, tl_eps 1e-15, tl_max_iters 5000, dt
0.004) over the number of timesteps given below. "config3": the two seeded
rectangles with conductivity 1/density (contrast 1e3), CG + strip-4 block
Jacobi for one timestep to tl_eps 1e-15 (capped at 8000 iterations, past the
FP64 rounding floor of the residual), the reference against which the refined
solves to a true residual of 1e-10 are measured. The oracle (oracle/tl_oracle.c, OpenMP)
needs minutes to an hour per run on a few cores, so the GPU tests compare
against these files instead of rerunning it.

Usage: python tests/golden/make_full_size.py [cg|chebyshev|ppcg|config3 ...]
"""
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

# solver name -> (deck keyword, timesteps)
RUNS = {"cg": ("tl_use_cg", 20), "chebyshev": ("tl_use_chebyshev", 20), "ppcg": ("tl_use_ppcg", 6),
        "config3": ("tl_use_cg", 1)}
N = 4096


def fixture_path(name):
    if name == "config3":
        return os.path.join(HERE, "config3_%d_cg.json" % N)
    return os.path.join(HERE, "config2_%d_%s.json" % (N, name))


def main(names):
    import pyoracle
    from tealeaf_opencl_amd import decks

    pyoracle.load()
    for name in names:
        solver, steps = RUNS[name]
        with tempfile.TemporaryDirectory() as d:
            if name == "config3":
                deck = decks.config3_deck(os.path.join(d, "c2.in"), precision="reference", n=N, steps=steps,
                                          eps=1e-15, max_iters=8000)
            else:
                deck = decks.config2_deck(os.path.join(d, "c2.in"), solver=solver, n=N, steps=steps)
            t0 = time.perf_counter()
            lo = pyoracle.run_deck(deck, os.path.join(d, "c2.log"))
            secs = time.perf_counter() - t0
        out = {"n": N, "solver": solver, "steps": steps, "seed": 1, "preconditioner": "jac_block",
               "oracle_threads": pyoracle.num_threads(), "oracle_seconds": secs,
               "summary": lo["summary"], "solve": lo["solve"]}
        with open(fixture_path(name), "w") as f:
            json.dump(out, f, indent=1)
        print("%s: %d steps in %.0f s" % (name, steps, secs), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(RUNS))
