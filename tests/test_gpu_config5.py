"""The large combined workload at its own size, bounded to a few minutes:
8192x8192 cells (67M; 537 MB per FP64 field), eight seeded rectangles with
densities 10^U[-1,2] (decks.config5_states), PPCG + strip-4 block Jacobi with
10 inner steps after 20 CG presteps, dt 0.004 -- one timestep each of the
reference FP64 PPCG (device-resident driver, its own exit test,
tea_solve.f90:231-267 with tea_leaf_ppcg_cl.cl:62-155) and of the refined
PPCG to a true residual of 1e-10 with FP64 and with FP32 inner fields; and
the density adjoint over 5 timesteps with one checkpoint interval of 5
(recompute of the whole interval), conductivity = density.

What holds at this size without a CPU run of the oracle (which would need
hours here): the volume is exactly 100 (every partial sum of 25 * 2^-24 is a
double); the mass equals dx dy times the numpy sum of the painted density
and does not change across the timestep; zero-flux walls make every column
of A sum to 1, so sum(u) = sum(u0) - sum(r) and the internal energy
sum(vol rho e1) moves only by the solve's residual; the refined solves'
true residuals, formed by the library's own residual and norm kernels, are
below 1e-10 of the initial ones; the mixed and FP64 refined summaries agree
within what their residuals allow; and the adjoint's directional derivative matches a central
difference of the GPU forward model."""
import numpy as np
import pytest

from tealeaf_opencl_amd import decks

pytestmark = pytest.mark.gpu

N = 8192
DT = 0.004


def painted_density_sum(n=N):
    """Sum and 2-norm over the interior of the density generate_chunk paints (numpy
    restatement of generate_chunk_cl.cl:3-65 for rectangles, with the dx/100
    shrink of read_input.f90:320-327)."""
    dx = 10.0 / float(np.float32(n))
    v = np.arange(n + 1) * dx  # vertices of the interior cells
    rho = np.full((n, n), 100.0)
    for s in decks.config5_states(1)[1:]:
        x0, x1 = s["xmin"] + dx / 100.0, s["xmax"] - dx / 100.0
        y0, y1 = s["ymin"] + dx / 100.0, s["ymax"] - dx / 100.0
        cx = (v[1:] >= x0) & (v[:-1] < x1)
        cy = (v[1:] >= y0) & (v[:-1] < y1)
        rho[np.ix_(cy, cx)] = s["density"]
    return float(np.sum(rho)), float(np.sqrt(np.sum(rho * rho))), dx


def test_config5_ppcg_one_timestep(hiplib, tmp_path):
    rho_sum, rho_norm, dx = painted_density_sum()
    mass_want = dx * dx * rho_sum
    # reference FP64 PPCG, device-resident driver
    deck = decks.config5_deck(str(tmp_path / "c5.in"), precision="reference", n=N, steps=1, summary_frequency=1)
    lg = hiplib.run_deck(deck, str(tmp_path / "c5.log"))
    s0, s1 = lg["summary"][0], lg["summary"][-1]
    assert s0[0] == 0 and s1[0] == 1
    assert s0[1] == 100.0 and s1[1] == 100.0
    assert s0[2] == s1[2] and abs(s0[2] - mass_want) <= 1e-13 * mass_want
    ie_ref = s1[4]
    # the reference's exit test (sqrt(r.z) < eps * initial) leaves a residual
    # whose sum moves the internal energy by ~1e-7 over 20 timesteps
    assert abs(s1[4] - s0[4]) <= 1e-6 * s0[4]
    print("reference PPCG: outer %d, inner %d, ie drift %.2e" % (lg["solve"][0][1], lg["solve"][0][4],
                                                                 abs(s1[4] - s0[4]) / s0[4]))

    out, r_norm = {}, {}
    for precision in (0, 1):
        c = hiplib.Chunk(N, N, 1, 3)
        try:
            c.initialise_chunk(0.0, 0.0, dx, dx)
            c.generate_chunk(*decks.state_arrays(decks.config5_states(1), N, N))
            c.set_field()
            before = c.field_summary()
            c.common_init(1, DT)
            c.calc_residual()
            rr0 = c.calc_2norm(1)
            res = c.solve_refined_ppcg(precision, tol=1e-10, max_iters=20000, presteps=20, inner_steps=10)
            assert res["converged"], res
            c.calc_residual()
            rel = np.sqrt(c.calc_2norm(1) / rr0)
            c.finalise()
            after = c.field_summary()
        finally:
            c.close()
        print("refined PPCG precision %d: %.2f s, outer %d, inner %d, updates %d, true residual %.2e "
              "(solver's %.2e)" % (precision, res["seconds"], res["iterations"], res["inner_iterations"],
                                   res["updates"], rel, res["rel_residual"]))
        assert rel <= 1e-10 and abs(rel - res["rel_residual"]) <= 1e-2 * 1e-10
        assert before[0] == 100.0 and after[0] == 100.0
        assert before[1] == after[1] and abs(after[1] - mass_want) <= 1e-13 * mass_want
        # sum(u) moves by -sum(r), |sum(r)| <= sqrt(N) |r|_2: the internal
        # energy sum(vol u) moves by at most vol * sqrt(N) * |r|_2 (plus the
        # rounding of u / rho * rho and of the sums)
        r_norm[precision] = rel * np.sqrt(rr0)
        drift = abs(after[2] - before[2])
        print("  internal energy drift %.2e, bound %.2e" % (drift, dx * dx * N * r_norm[precision]))
        assert drift <= dx * dx * N * r_norm[precision] + 1e-12 * before[2]
        out[precision] = np.array(after)
    d = np.abs(out[1] - out[0])
    print("mixed vs FP64 refined (relative): volume %.1e mass %.1e ie %.2e temperature %.2e"
          % tuple(d / np.abs(out[0])))
    assert d[0] == 0 and d[1] == 0
    # both drifts bound the internal energies' difference; the eigenvalues of
    # A are >= 1, so |u_a - u_b|_2 <= |r_a|_2 + |r_b|_2 and the temperatures
    # sum(vol rho u) differ by at most vol |rho|_2 (|r_a|_2 + |r_b|_2)
    r_sum = r_norm[0] + r_norm[1]
    assert d[2] <= dx * dx * N * r_sum + 1e-12 * out[0][2]
    assert d[3] <= dx * dx * rho_norm * r_sum + 1e-12 * out[0][3]
    # the refined solves and the reference PPCG solve one timestep: the
    # internal energies agree to what the reference's exit test leaves
    assert abs(out[0][2] - ie_ref) <= 1e-6 * ie_ref


def test_config5_adjoint_checkpointed(hiplib):
    """The density adjoint at 8192^2 over 5 timesteps, checkpoint every 5:
    one checkpoint, every forward state of the interval recomputed. Device
    work space is the checkpoint, five recomputed states and four vectors
    (u_obs, the adjoint seed, e0, the gradient); the directional derivative
    along rho * N(0,1) over every cell matches the central difference with
    eps = 1e-6 (measured at 20 timesteps: 1.6e-5, the forward solves' FP64
    floor; held to 1e-4 here)."""
    from tealeaf_opencl_amd import adjoint
    nsteps, tol, coef = 5, 1e-12, 1
    dx = 10.0 / float(np.float32(N))
    c = hiplib.Chunk(N, N, 1, 3)
    try:
        c.initialise_chunk(0.0, 0.0, dx, dx)
        c.generate_chunk(*decks.state_arrays(decks.config5_states(2), N, N))
        c.set_field()
        e0 = c.get_field("energy1")
        u_obs = adjoint.forward_u(c, coef, DT, nsteps, tol)
        c.generate_chunk(*decks.state_arrays(decks.config5_states(1), N, N))
        c.set_field()
        rho, e0 = c.get_field("density"), c.get_field("energy1")
        L, grad, info = c.adjoint_density(coef, DT, nsteps, u_obs, tol=tol, checkpoint_every=5)
        field = info["work_bytes"] / 10.0
        assert info["checkpoints"] == 1
        assert field == int(field) and (N + 2) ** 2 * 8 <= field <= 1.01 * (N + 2) ** 2 * 8
        # the chunk's 23 fields and the FP64 refined solve's increment field
        assert info["chunk_bytes"] == 24 * field
        assert info["forward_iterations"] > 0 and info["recompute_iterations"] > 0
        assert info["adjoint_iterations"] > 0
        print("adjoint 8192^2 x %d: forward %.2f s, adjoint %.2f s (%.2fx), device %.1f GB, stalled %d"
              % (nsteps, info["forward_s"], info["adjoint_s"], info["adjoint_s"] / info["forward_s"],
                 (info["chunk_bytes"] + info["work_bytes"]) / 1e9, info["stalled_solves"]))
        halo = np.ones(grad.shape, dtype=bool)
        halo[1:-1, 1:-1] = False
        assert np.all(grad[halo] == 0.0) and np.all(np.isfinite(grad))
        v = np.zeros_like(rho)
        v[1:-1, 1:-1] = np.random.default_rng(3).standard_normal((N, N)) * rho[1:-1, 1:-1]
        dd = float(np.sum(grad * v))
        eps = 1e-6
        vals = []
        for s in (1.0, -1.0):
            c.put_field("density", rho + s * eps * v)
            c.put_field("energy1", e0)
            vals.append(c.forward_loss(coef, DT, nsteps, u_obs, tol=tol))
        fd = (vals[0] - vals[1]) / (2 * eps)
        err = abs(fd - dd) / abs(dd)
        print("directional derivative: adjoint %.10e, central difference %.10e, rel %.2e" % (dd, fd, err))
        assert err <= 1e-4
    finally:
        c.close()
