"""Writes the CPU reference adjoint of the 1024x1024 adjoint workload as a
fixture: per-cell density decks.config4_density(1024, seed 1), energy from the
two seeded rectangles (seed 1), conductivity = density, strip-4 block Jacobi,
dt 0.004, 5 timesteps; u_obs from the same forward model on the seed-2
density. Every forward and adjoint solve is the oracle's FP64 CG
(oracle/tl_oracle.c through tests/_adjoint_ref.py, CG run to a recursive
residual of 1e-14 of the initial one); the face gradients are scattered to
both cells of each face in numpy.

Stored: the loss L, max |dL/drho|, the sums of dL/drho and |dL/drho| over the
interior, and dL/drho at 32 sensitive cells (those the finite-difference test
draws: seed 2024 among the cells with |rho dL/drho| >= 1e-2 of its maximum)
and at 2000 seeded interior cells. The full gradient (8 MB) is not kept.

Usage: python tests/golden/make_config4_adjoint.py   (about 10-20 minutes on 8 cores)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

N, SEED, OBS_SEED, NSTEPS, DT, COEF, PREC, TOL = 1024, 1, 2, 5, 0.004, 1, 3, 1e-14
PATH = os.path.join(HERE, "config4_%d_adjoint.json" % N)


def sensitive_cells(rho, grad, count, seed, frac=1e-2):
    """The cell choice of the GPU finite-difference test (tests/test_gpu_adjoint.py)."""
    s = np.abs(rho * grad)
    ks, js = np.nonzero(s >= frac * s.max())
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(ks), size=min(count, len(ks)), replace=False)
    return [(int(ks[i]), int(js[i])) for i in pick]


def main():
    import pyoracle
    import _adjoint_ref as ref
    from tealeaf_opencl_amd import decks

    pyoracle.load()
    t0 = time.perf_counter()
    o = pyoracle.OracleChunk(N, N, 1, PREC)
    dx = 10.0 / float(np.float32(N))
    o.initialise_chunk(0.0, 0.0, dx, dx)
    o.generate_chunk(*decks.state_arrays(decks.seeded_states(SEED, 2, False), N, N))
    o.set_field()
    e0 = o.get_field("energy1").copy()
    rho = decks.config4_density(N, SEED, 1)
    rho_obs = decks.config4_density(N, OBS_SEED, 1)
    o.put_field("density", rho_obs)
    o.put_field("energy1", e0)
    us, _ = ref.forward(o, COEF, DT, NSTEPS, TOL)
    u_obs = us[-1].copy()
    print("u_obs: %.0f s" % (time.perf_counter() - t0), flush=True)
    o.put_field("density", rho)
    o.put_field("energy1", e0)
    L, grad = ref.adjoint_grad(o, COEF, DT, NSTEPS, u_obs, tol=TOL)
    secs = time.perf_counter() - t0
    inner = grad[1:-1, 1:-1]
    cells = sensitive_cells(rho, grad, 32, 2024)
    rng = np.random.default_rng(77)
    sample = [(int(k), int(j)) for k, j in zip(rng.integers(1, N + 1, 2000), rng.integers(1, N + 1, 2000))]
    out = {"n": N, "seed": SEED, "obs_seed": OBS_SEED, "nsteps": NSTEPS, "dt": DT, "coefficient": COEF,
           "preconditioner": PREC, "cg_tol": TOL, "oracle_threads": pyoracle.num_threads(),
           "oracle_seconds": secs, "loss": L, "grad_max_abs": float(np.max(np.abs(inner))),
           "grad_sum": float(np.sum(inner)), "grad_abs_sum": float(np.sum(np.abs(inner))),
           "u_obs_sum": float(np.sum(u_obs[1:-1, 1:-1])),
           "sensitive": [[k, j, float(grad[k, j])] for k, j in cells],
           "sample": [[k, j, float(grad[k, j])] for k, j in sample]}
    with open(PATH, "w") as f:
        json.dump(out, f, indent=0)
    print("L = %.17g, %.0f s" % (L, secs), flush=True)


if __name__ == "__main__":
    main()
