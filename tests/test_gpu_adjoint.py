"""Density adjoint on the GPU (tl_adjoint_density): against the oracle-based
reference gradient (tests/_adjoint_ref.py, itself pinned by finite
differences on the CPU), against central finite differences of the GPU
forward model (h = 1e-6 relative, 32 seeded cells), through the
torch.autograd wrapper, and across checkpoint intervals."""
import numpy as np
import pytest

import _adjoint_ref as ref

pytestmark = pytest.mark.gpu


def random_problem(x, y, seed):
    rng = np.random.default_rng(seed)
    rho = np.zeros((y + 2, x + 2))
    rho[1:-1, 1:-1] = 10.0 ** rng.uniform(-1.0, 2.0, (y, x))
    e0 = np.zeros_like(rho)
    e0[1:-1, 1:-1] = rng.uniform(0.5, 2.0, (y, x))
    u_obs = np.zeros_like(rho)
    u_obs[1:-1, 1:-1] = rng.uniform(0.0, 50.0, (y, x))
    return rho, e0, u_obs


def gpu_chunk(hiplib, rho, e0, x, y, prec):
    c = hiplib.Chunk(x, y, 1, prec)
    c.initialise_chunk(0.0, 0.0, 10.0 / float(np.float32(x)), 10.0 / float(np.float32(y)))
    c.put_field("density", rho)
    c.put_field("energy1", e0)
    return c


@pytest.mark.parametrize("coef", [1, 2])
@pytest.mark.parametrize("prec", [1, 2, 3])
@pytest.mark.parametrize("x,y", [(96, 64), (67, 13)])
def test_against_oracle_reference(hiplib, oracle, x, y, prec, coef):
    nsteps, dt = 3, 0.004
    rho, e0, u_obs = random_problem(x, y, 11)
    g = gpu_chunk(hiplib, rho, e0, x, y, prec)
    L, grad, info = g.adjoint_density(coef, dt, nsteps, u_obs, tol=1e-13, checkpoint_every=2)
    o = ref.oracle_chunk(oracle, rho, e0, x, y, prec=prec)
    L_ref, grad_ref = ref.adjoint_grad(o, coef, dt, nsteps, u_obs, tol=1e-15)
    assert abs(L - L_ref) <= 1e-10 * L_ref
    scale = np.max(np.abs(grad_ref))
    assert np.max(np.abs(grad - grad_ref)) <= 1e-8 * scale
    halo = np.ones(grad.shape, dtype=bool)
    halo[1:-1, 1:-1] = False
    assert np.all(grad[halo] == 0.0)
    assert info["checkpoints"] == 2 and info["forward_iterations"] > 0
    assert info["recompute_iterations"] > 0 and info["adjoint_iterations"] > 0
    # the final forward energy is left in energy1
    e_gpu, e_ref = g.interior(g.get_field("energy1")), o.interior(o.get_field("energy1"))
    assert np.max(np.abs(e_gpu - e_ref)) <= 1e-10 * np.max(np.abs(e_ref))


def sensitive_cells(rho, grad, count, seed, frac=1e-2):
    """`count` seeded interior cells among those whose density moves the loss
    by at least `frac` of the largest |rho dL/drho|. Over 5 steps of dt 0.004
    heat moves a few cells, so most cells' sensitivity is below what a
    central difference of a loss of ~1e7 can resolve in FP64."""
    s = np.abs(rho * grad)
    ks, js = np.nonzero(s >= frac * s.max())
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(ks), size=min(count, len(ks)), replace=False)
    return [(int(ks[i]), int(js[i])) for i in pick]


def test_finite_differences_config4(hiplib):
    """The adjoint workload's layout at 128x128: 32 seeded cells (among the
    sensitive ones), central differences with h = 1e-6 relative, 5
    timesteps."""
    from tealeaf_opencl_amd import adjoint
    n, nsteps, dt, coef = 128, 5, 0.004, 1
    c = hiplib.Chunk(n, n, 1, 3)
    rho, e0, u_obs = adjoint.setup_problem(c, seed=1, obs_seed=2, dt=dt, nsteps=nsteps, tol=1e-14)
    L, grad, _ = c.adjoint_density(coef, dt, nsteps, u_obs, tol=1e-14, checkpoint_every=2)
    cells = sensitive_cells(rho, grad, 32, 2024)
    assert len(cells) == 32
    worst = 0.0
    for (k, j) in cells:
        h = 1e-6 * rho[k, j]
        vals = []
        for s in (1.0, -1.0):
            r2 = rho.copy()
            r2[k, j] += s * h
            c.put_field("density", r2)
            c.put_field("energy1", e0)
            vals.append(c.forward_loss(coef, dt, nsteps, u_obs, tol=1e-14))
        fd = (vals[0] - vals[1]) / (2.0 * h)
        worst = max(worst, abs(fd - grad[k, j]) / abs(grad[k, j]))
    assert worst <= 1e-6, worst


def test_finite_differences_config4_full_size(hiplib):
    """The adjoint workload at its own size, 1024x1024, 5 timesteps: 32
    seeded sensitive cells, central differences with h = 1e-4 relative. At
    this size a central difference resolves the loss only to the FP64
    forward solves' rounding floor: with h = 1e-6 the error against the
    adjoint is that floor over h (median 6e-6, worst 8e-5 of the 32 cells),
    and it falls as 1/h (h = 1e-5: 9e-7 / 7e-6); at h = 1e-4 the floor's
    share is ~1e-7 median / ~1e-6 worst, and the O(h^2) truncation is below
    1e-8 (h = 1e-3 gives 3e-8 / 7e-7). Which rounding the floor takes moves
    with the order of the solves' sums, so the bounds are three times those
    figures. (The exact check of the adjoint at this size is the comparison
    with the CPU reference adjoint below.)"""
    from tealeaf_opencl_amd import adjoint
    n, nsteps, dt, coef = 1024, 5, 0.004, 1
    c = hiplib.Chunk(n, n, 1, 3)
    rho, e0, u_obs = adjoint.setup_problem(c, seed=1, obs_seed=2, dt=dt, nsteps=nsteps, tol=1e-14)
    L, grad, _ = c.adjoint_density(coef, dt, nsteps, u_obs, tol=1e-14, checkpoint_every=2)
    cells = sensitive_cells(rho, grad, 32, 2024)
    assert len(cells) == 32
    errs = []
    for (k, j) in cells:
        h = 1e-4 * rho[k, j]
        vals = []
        for s in (1.0, -1.0):
            r2 = rho.copy()
            r2[k, j] += s * h
            c.put_field("density", r2)
            c.put_field("energy1", e0)
            vals.append(c.forward_loss(coef, dt, nsteps, u_obs, tol=1e-14))
        fd = (vals[0] - vals[1]) / (2.0 * h)
        errs.append(abs(fd - grad[k, j]) / abs(grad[k, j]))
    print("h = 1e-4: median %.2e, worst %.2e" % (np.median(errs), max(errs)))
    assert np.median(errs) <= 3e-7 and max(errs) <= 3e-6, (np.median(errs), max(errs))


def test_config4_full_size_vs_cpu_reference(hiplib):
    """The adjoint workload at 1024x1024 over 5 timesteps against the CPU
    reference adjoint of the same problem (fixture
    tests/golden/config4_1024_adjoint.json, written by
    tests/golden/make_config4_adjoint.py: every solve the oracle's FP64 CG, the
    face terms scattered in numpy): the loss to 1e-10 relative, dL/drho at the
    32 sensitive cells and 2000 seeded cells to 1e-8 of the gradient's
    largest component, and the interior sum of dL/drho. The GPU side forms
    its own u_obs (refined FP64 solves) from the seed-2 density."""
    import json
    import os
    from tealeaf_opencl_amd import adjoint
    path = os.path.join(os.path.dirname(__file__), "golden", "config4_1024_adjoint.json")
    with open(path) as f:
        fx = json.load(f)
    n, nsteps, dt, coef = fx["n"], fx["nsteps"], fx["dt"], fx["coefficient"]
    c = hiplib.Chunk(n, n, 1, fx["preconditioner"])
    rho, e0, u_obs = adjoint.setup_problem(c, seed=fx["seed"], obs_seed=fx["obs_seed"], dt=dt, nsteps=nsteps,
                                           tol=1e-14)
    assert abs(float(np.sum(u_obs[1:-1, 1:-1])) - fx["u_obs_sum"]) <= 1e-12 * abs(fx["u_obs_sum"])
    L, grad, _ = c.adjoint_density(coef, dt, nsteps, u_obs, tol=1e-14, checkpoint_every=2)
    scale = fx["grad_max_abs"]
    assert abs(np.max(np.abs(grad[1:-1, 1:-1])) - scale) <= 1e-8 * scale
    errs = [abs(grad[k, j] - v) / scale for k, j, v in fx["sensitive"] + fx["sample"]]
    rel_sens = [abs(grad[k, j] - v) / abs(v) for k, j, v in fx["sensitive"]]
    sum_err = abs(float(np.sum(grad[1:-1, 1:-1])) - fx["grad_sum"]) / fx["grad_abs_sum"]
    print("config4 vs CPU reference: L %.2e, components %.2e (of max), sensitive %.2e (own), sum %.2e"
          % (abs(L - fx["loss"]) / fx["loss"], max(errs), max(rel_sens), sum_err))
    assert abs(L - fx["loss"]) <= 1e-10 * fx["loss"]
    assert max(errs) <= 1e-8
    assert max(rel_sens) <= 1e-6
    assert sum_err <= 1e-8


def test_torch_wrapper(hiplib):
    torch = pytest.importorskip("torch")
    from tealeaf_opencl_amd import adjoint
    n = 64
    c = hiplib.Chunk(n, n, 1, 2)
    rho, e0, u_obs = adjoint.setup_problem(c, nsteps=3, tol=1e-13)
    t = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
    loss = adjoint.density_loss(t, c, e0, u_obs, nsteps=3, tol=1e-13, checkpoint_every=1)
    loss.backward()
    c.put_field("density", rho)
    c.put_field("energy1", e0)
    L, grad, _ = c.adjoint_density(1, 0.004, 3, u_obs, tol=1e-13, checkpoint_every=1)
    assert float(loss) == L
    assert np.array_equal(t.grad.numpy(), grad)
    # directional derivative against central differences along a random
    # interior direction
    rng = np.random.default_rng(3)
    v = np.zeros_like(rho)
    v[1:-1, 1:-1] = rng.standard_normal((n, n)) * rho[1:-1, 1:-1]
    eps = 1e-6
    vals = []
    for s in (1.0, -1.0):
        c.put_field("density", rho + s * eps * v)
        c.put_field("energy1", e0)
        vals.append(c.forward_loss(1, 0.004, 3, u_obs, tol=1e-14))
    fd = (vals[0] - vals[1]) / (2 * eps)
    dd = float(np.sum(grad * v))
    assert abs(fd - dd) <= 1e-6 * abs(dd)


def test_checkpoint_intervals_agree(hiplib):
    x, y, nsteps = 80, 72, 5
    rho, e0, u_obs = random_problem(x, y, 4)
    grads = []
    for every in (1, 2, 5, 7):
        c = gpu_chunk(hiplib, rho, e0, x, y, 3)
        L, grad, info = c.adjoint_density(1, 0.004, nsteps, u_obs, tol=1e-13, checkpoint_every=every)
        assert info["checkpoints"] == -(-nsteps // min(every, nsteps))
        grads.append(grad)
    scale = np.max(np.abs(grads[0]))
    for gr in grads[1:]:
        assert np.max(np.abs(gr - grads[0])) <= 1e-10 * scale


def test_bad_arguments(hiplib):
    rho, e0, u_obs = random_problem(16, 16, 1)
    c = gpu_chunk(hiplib, rho, e0, 16, 16, 1)
    for kw in ({"nsteps": 0}, {"checkpoint_every": 0}, {"coefficient": 3}, {"tol": 0.0}):
        args = dict(coefficient=1, dt=0.004, nsteps=2, u_obs=u_obs, tol=1e-12, checkpoint_every=1)
        args.update(kw)
        with pytest.raises(hiplib.TealeafError):
            c.adjoint_density(**args)
    with pytest.raises(ValueError):
        c.adjoint_density(1, 0.004, 2, u_obs[:-1], tol=1e-12)


def test_progress_callback(hiplib):
    x, y, nsteps = 48, 40, 4
    rho, e0, u_obs = random_problem(x, y, 9)
    c = gpu_chunk(hiplib, rho, e0, x, y, 2)
    seen = []
    c.set_progress(lambda phase, step, its: seen.append((phase, step, its)))
    c.adjoint_density(1, 0.004, nsteps, u_obs, tol=1e-12, checkpoint_every=2)
    phases = [p for p, _, _ in seen]
    assert phases.count("forward") == nsteps
    assert phases.count("recompute") == nsteps and phases.count("adjoint") == nsteps
    assert [s for p, s, _ in seen if p == "adjoint"] == [4, 3, 2, 1]
    assert all(its > 0 for _, _, its in seen)
    c.set_progress(None)
    seen.clear()
    c.put_field("energy1", e0)
    c.forward_loss(1, 0.004, 2, u_obs)
    assert not seen
