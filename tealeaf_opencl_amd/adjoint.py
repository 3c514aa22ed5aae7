"""Density gradients of the implicit timesteps: the adjoint workload's set-up
and a thin torch.autograd wrapper over Chunk.adjoint_density (the C-ABI entry
point tl_adjoint_density), used for gradient tests.

Workload (1024x1024 by default): per-cell density from decks.config4_density,
energy from the two seeded rectangles of the 512x512 CG deck (background
1e-4, rectangles 25), conductivity = density, dt 0.004, 5 FP64 timesteps;
u_obs is the final u of the same forward model run on a second seeded
density field; L = sum over the interior of (u_T - u_obs)^2.
"""
import numpy as np

from . import (COEF_CONDUCTIVITY, PRECISION_FP64, Chunk, decks)


def forward_u(chunk, coefficient, dt, nsteps, tol=1e-12, progress=None):
    """u after `nsteps` timesteps from the chunk's energy1 (tl_common_init,
    tl_solve_refined in FP64, tl_finalise per step); a solve that stops at its
    rounding floor above tol is accepted, as in the library's adjoint.
    progress(phase, step, iterations) is called after every solve."""
    for step in range(1, nsteps + 1):
        chunk.common_init(coefficient, dt)
        res = chunk.solve_refined(PRECISION_FP64, tol=tol, max_iters=20000)
        if not (res["converged"] or res["stalled"]):
            raise RuntimeError("forward solve did not converge: %r" % (res,))
        chunk.finalise()
        if progress:
            progress("forward", step, res["iterations"])
    return chunk.get_field("u")


def setup_problem(chunk, seed=1, obs_seed=2, coefficient=COEF_CONDUCTIVITY, dt=0.004, nsteps=5, tol=1e-12):
    """Initialise `chunk` for the adjoint workload. Returns (rho, e0, u_obs):
    the density to differentiate at, the initial energy and the observation.
    On return the chunk holds rho and e0."""
    if chunk.x_max != chunk.y_max:
        raise ValueError("the adjoint workload is square")
    n = chunk.x_max
    dx = 10.0 / float(np.float32(n))
    chunk.initialise_chunk(0.0, 0.0, dx, dx)
    chunk.generate_chunk(*decks.state_arrays(decks.seeded_states(seed, 2, False), n, n))
    chunk.set_field()
    e0 = chunk.get_field("energy1")
    rho = decks.config4_density(n, seed, chunk.hd)
    rho_obs = decks.config4_density(n, obs_seed, chunk.hd)
    chunk.put_field("density", rho_obs)
    chunk.put_field("energy1", e0)
    u_obs = forward_u(chunk, coefficient, dt, nsteps, tol)
    chunk.put_field("density", rho)
    chunk.put_field("energy1", e0)
    return rho, e0, u_obs


def density_loss(density, chunk, energy, u_obs, coefficient=COEF_CONDUCTIVITY, dt=0.004, nsteps=5, tol=1e-12,
                 checkpoint_every=5):
    """L(density) as a torch scalar whose backward yields dL/d density from
    the library's adjoint. `density` is a float64 tensor in the layout of
    Chunk.get_field; `energy` the initial energy1 (numpy, same layout)."""
    return _function().apply(density, chunk, energy, u_obs, coefficient, dt, nsteps, tol, checkpoint_every)


_FN = None


def _function():
    global _FN
    if _FN is None:
        import torch

        class DensityLoss(torch.autograd.Function):
            @staticmethod
            def forward(ctx, density, chunk, energy, u_obs, coefficient, dt, nsteps, tol, checkpoint_every):
                rho = density.detach().to("cpu", torch.float64).numpy()
                chunk.put_field("density", rho)
                chunk.put_field("energy1", energy)
                loss, grad, _ = chunk.adjoint_density(coefficient, dt, nsteps, u_obs, tol=tol,
                                                      checkpoint_every=checkpoint_every)
                ctx.save_for_backward(torch.from_numpy(grad).to(density.device))
                return density.new_tensor(loss)

            @staticmethod
            def backward(ctx, grad_out):
                (grad,) = ctx.saved_tensors
                return (grad_out * grad,) + (None,) * 8

        _FN = DensityLoss
    return _FN


__all__ = ["forward_u", "setup_problem", "density_loss", "Chunk"]
