"""CPU reference for the density adjoint, built on the oracle (test
infrastructure only): the forward timesteps and every adjoint solve use the
oracle's FP64 CG (oracle/tl_oracle.c, tea_leaf_cg_cl.cl restated); the face
gradients are scattered to the two cells of each face here in numpy, an
arrangement independent of the GPU's per-cell gather."""
import numpy as np

from _pairs import oracle_cg


def oracle_chunk(oracle, rho, e0, x, y, hd=1, prec=1):
    o = oracle.OracleChunk(x, y, hd, prec)
    o.initialise_chunk(0.0, 0.0, 10.0 / float(np.float32(x)), 10.0 / float(np.float32(y)))
    o.put_field("density", rho)
    o.put_field("energy1", e0)
    return o


def forward(o, coef, dt, nsteps, tol):
    """u^0 .. u^T of the timesteps from the oracle's energy1, and (rx, ry)."""
    us = []
    rr = None
    for _ in range(nsteps):
        rr = o.common_init(coef, dt)
        us.append(o.get_field("u0").copy())
        oracle_cg(o, tol)
        o.finalise()
    us.append(o.get_field("u").copy())
    return us, rr


def loss(o, coef, dt, nsteps, u_obs, tol=1e-15):
    us, _ = forward(o, coef, dt, nsteps, tol)
    d = o.interior(us[-1] - u_obs)
    return float(np.sum(d * d))


def dk_drho(rho, r, coef):
    """dK/drho of K = r (w_a + w_b) / (2 w_a w_b) w.r.t. each side's density,
    w = rho (coef 1) or 1/rho (coef 2); zero where rho is zero (halo)."""
    safe = np.where(rho != 0.0, rho, 1.0)
    if coef == 1:
        return np.where(rho != 0.0, -r / (2.0 * safe * safe), 0.0)
    return np.where(rho != 0.0, 0.5 * r, 0.0)


def adjoint_grad(o, coef, dt, nsteps, u_obs, tol=1e-15):
    """(L, dL/d density) on the oracle chunk (state restored from its energy1
    and density on entry)."""
    e0 = o.get_field("energy1").copy()
    rho = o.get_field("density").copy()
    us, (rx, ry) = forward(o, coef, dt, nsteps, tol)
    kx, ky = o.get_field("Kx"), o.get_field("Ky")
    free_x = kx[:, 1:] != 0.0          # face between columns j-1 and j, j >= 1
    free_y = ky[1:, :] != 0.0          # face between rows k-1 and k, k >= 1
    dxr, dxl = dk_drho(rho[:, 1:], rx, coef), dk_drho(rho[:, :-1], rx, coef)
    dyt, dyb = dk_drho(rho[1:, :], ry, coef), dk_drho(rho[:-1, :], ry, coef)
    g = np.zeros_like(rho)
    o.interior(g)[:] = 2.0 * o.interior(us[-1] - u_obs)
    L = float(np.sum(o.interior(us[-1] - u_obs) ** 2))
    grad = np.zeros_like(rho)
    for n in range(nsteps - 1, -1, -1):
        o.put_field("u0", g)
        o.put_field("u", np.zeros_like(g))
        lam = oracle_cg(o, tol).copy()
        u1 = us[n + 1]
        gx = np.where(free_x, -(lam[:, 1:] - lam[:, :-1]) * (u1[:, 1:] - u1[:, :-1]), 0.0)
        gy = np.where(free_y, -(lam[1:, :] - lam[:-1, :]) * (u1[1:, :] - u1[:-1, :]), 0.0)
        grad[:, 1:] += gx * dxr
        grad[:, :-1] += gx * dxl
        grad[1:, :] += gy * dyt
        grad[:-1, :] += gy * dyb
        g = lam
    grad += g * e0
    return L, grad
