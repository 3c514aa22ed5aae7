"""Pins of the CPU oracle (oracle/tl_oracle.c) that do not use the GPU.

What pins what:
* whole path (deck reader, generate, driver, Jacobi/PPCG solvers, field
  summary): the reference's known-answer temp values, field_summary.f90:90-95,
  tests 1 and 6 (test 2 is run on the GPU and by tests/golden/make_kat2.sh);
* face coefficients: closed forms of both coefficient kinds (conductivity =
  density and = 1/density) on every face of a random density, the walls
  exactly zero;
* reflective update_halo: a mirror index map of each external face's halo band
  (corners included), its symmetry, and that an off-by-one mirror differs;
* operator: A 1 = 1 with Neumann walls, symmetry <Ax,y> = <x,Ay>;
* solvers: dense numpy solves of A u = u0 on small grids (CG none/diag/block,
  Jacobi, and whole-deck Chebyshev/PPCG via the field summary);
* block Jacobi: the 4x4 tridiagonal strip systems hold exactly;
* tqli / eigenvalue estimate: LAPACK eigvalsh and a Lanczos run on a
  matrix with known spectrum;
* Chebyshev coefficients: the closed form rho_n = T_n(s)/T_{n+1}(s).
"""
import math
import os

import numpy as np
import pytest

from _pairs import setup
from tealeaf_opencl_amd import decks


def _kat(oracle, decks_dir, tmp_path, n):
    return oracle.run_deck(os.path.join(decks_dir, "kat%d.in" % n), tmp_path / ("kat%d.log" % n))


def test_kat1_jacobi(oracle, decks_dir, tmp_path):
    lg = _kat(oracle, decks_dir, tmp_path, 1)
    assert lg["qa"][2] == "PASSED"
    assert lg["qa"][1] < 1e-9  # percent
    assert abs(lg["summary"][-1][5] - 157.550841832793) < 1e-11 * 157.55


def test_kat6_ppcg(oracle, decks_dir, tmp_path):
    lg = _kat(oracle, decks_dir, tmp_path, 6)
    assert lg["qa"][2] == "PASSED"
    assert abs(lg["summary"][-1][5] - 103.88639125996923) < 1e-10 * 103.9
    # zero-flux walls: ie = sum vol*u and the mass are conserved across steps
    ie = [s[4] for s in lg["summary"]]
    mass = [s[2] for s in lg["summary"]]
    assert max(abs(x - ie[0]) for x in ie) < 1e-12 * ie[0]
    assert max(abs(x - mass[0]) for x in mass) < 1e-12 * mass[0]
    # PPCG switches at n=2 on every step after the first (tea_solve.f90:155-157)
    assert all(s[1] == 2 for s in lg["switch"][1:])


def _dense_operator(o):
    """Assemble A from the oracle's Kx/Ky on the interior (zero walls)."""
    Kx, Ky = o.get_field("Kx"), o.get_field("Ky")
    hd, nx, ny = o.hd, o.x_max, o.y_max
    N = nx * ny
    A = np.zeros((N, N))
    idx = lambda j, k: (k - hd) * nx + (j - hd)
    for k in range(hd, hd + ny):
        for j in range(hd, hd + nx):
            i = idx(j, k)
            A[i, i] = 1.0 + Kx[k, j] + Kx[k, j + 1] + Ky[k, j] + Ky[k + 1, j]
            if j > hd:
                A[i, idx(j - 1, k)] = -Kx[k, j]
            if j < hd + nx - 1:
                A[i, idx(j + 1, k)] = -Kx[k, j + 1]
            if k > hd:
                A[i, idx(j, k - 1)] = -Ky[k, j]
            if k < hd + ny - 1:
                A[i, idx(j, k + 1)] = -Ky[k + 1, j]
    return A


def test_face_coefficients_closed_form(oracle):
    o = oracle.OracleChunk(40, 40, 1, 1)
    rx, ry = setup(o, seed=1, circle=False)
    rho = o.get_field("density")
    Kx = o.get_field("Kx")
    dx = 10.0 / 40
    assert rx == 0.004 / (dx * dx)
    for k in range(2, 40):
        for j in range(2, 40):
            dl, dc = rho[k, j - 1], rho[k, j]
            # conductivity = density: K = rx * mean(1/rho) over the two cells
            want = rx * 0.5 * (1.0 / dl + 1.0 / dc)
            assert abs(Kx[k, j] - want) <= 1e-14 * want
    # walls are zero-flux
    assert np.all(Kx[:, 1] == 0) and np.all(Kx[:, 41] == 0)


@pytest.mark.parametrize("coef", [1, 2])
@pytest.mark.parametrize("hd", [1, 2])
def test_face_coefficients_every_face(oracle, coef, hd):
    """tea_leaf_init_common + tea_leaf_zero_boundary
    (tea_leaf_common_cl.cl:56-125) on a log-uniform random density, both
    coefficient kinds, against their closed forms: conductivity = density
    gives K = rx (1/rho_l + 1/rho_c) / 2, conductivity = 1/density gives
    K = rx (rho_l + rho_c) / 2. With all four walls external, every face that
    is not between two interior cells is exactly 0."""
    nx, ny = 23, 17
    o = oracle.OracleChunk(nx, ny, hd, 1)
    setup(o, seed=2)
    rng = np.random.default_rng(11)
    rho = 10.0 ** rng.uniform(-1, 2, o.shape)
    rho_in = np.zeros(o.shape)
    o.interior(rho_in)[:] = o.interior(rho)
    o.put_field("density", rho_in)
    rx, ry = o.common_init(coef, 0.004)
    Kx, Ky = o.get_field("Kx"), o.get_field("Ky")

    def closed(a, b, r):
        return r * 0.5 * (1.0 / a + 1.0 / b) if coef == 1 else r * 0.5 * (a + b)

    want_x = np.zeros(o.shape)
    want_y = np.zeros(o.shape)
    rows, cols = slice(hd, hd + ny), slice(hd, hd + nx)
    # x faces between interior columns j-1 and j, j = hd+1 .. hd+nx-1
    want_x[rows, hd + 1:hd + nx] = closed(rho_in[rows, hd:hd + nx - 1], rho_in[rows, hd + 1:hd + nx], rx)
    want_y[hd + 1:hd + ny, cols] = closed(rho_in[hd:hd + ny - 1, cols], rho_in[hd + 1:hd + ny, cols], ry)
    mask_x, mask_y = want_x != 0, want_y != 0
    assert np.all(Kx[~mask_x] == 0.0) and np.all(Ky[~mask_y] == 0.0)
    assert np.max(np.abs(Kx[mask_x] - want_x[mask_x]) / want_x[mask_x]) <= 2e-15
    assert np.max(np.abs(Ky[mask_y] - want_y[mask_y]) / want_y[mask_y]) <= 2e-15
    # the two kinds differ on every interior face of a non-uniform field
    other = (rx * 0.5 * (rho_in[rows, hd:hd + nx - 1] + rho_in[rows, hd + 1:hd + nx]) if coef == 1 else
             rx * 0.5 * (1.0 / rho_in[rows, hd:hd + nx - 1] + 1.0 / rho_in[rows, hd + 1:hd + nx]))
    assert np.all(np.abs(Kx[rows, hd + 1:hd + nx] - other) > 1e-6 * other)


def _reflected(a0, hd, nx, ny, depth, ext):
    """The reflective fill stated as an index map, apart from the kernels'
    offset arithmetic (update_halo_cl.cl:3-69, extents ocl_kernels.cpp:356-412):
    a cell of the halo band of an external face, within `depth` of the
    interior and spanning the corner squares of that depth, takes the value of
    its mirror image across that wall (and across the other wall too in a
    corner whose other face is also external); every other cell keeps its
    value. ext = (left, right, bottom, top)."""
    ny2, nx2 = a0.shape
    left, right, bottom, top = ext

    def mirror(i, lo, n, lo_ext, hi_ext):
        # reflection across the wall between lo-1|lo or lo+n-1|lo+n
        if i < lo and lo_ext:
            return 2 * lo - 1 - i
        if i >= lo + n and hi_ext:
            return 2 * (lo + n) - 1 - i
        return i

    out = a0.copy()
    band_x = lambda j: hd - depth <= j < hd + nx + depth
    band_y = lambda k: hd - depth <= k < hd + ny + depth
    for k in range(ny2):
        for j in range(nx2):
            in_b = bottom and hd - depth <= k < hd and band_x(j)
            in_t = top and hd + ny <= k < hd + ny + depth and band_x(j)
            in_l = left and hd - depth <= j < hd and band_y(k)
            in_r = right and hd + nx <= j < hd + nx + depth and band_y(k)
            if in_b or in_t or in_l or in_r:
                out[k, j] = a0[mirror(k, hd, ny, bottom, top), mirror(j, hd, nx, left, right)]
    return out


@pytest.mark.parametrize("hd,depth", [(1, 1), (2, 1), (2, 2), (3, 1), (3, 2), (3, 3)])
@pytest.mark.parametrize("ext", [(1, 1, 1, 1), (1, 0, 1, 0), (0, 1, 0, 1), (0, 0, 1, 0)])
def test_update_halo_reflective_mirror(oracle, hd, depth, ext):
    """The oracle's update_halo against the mirror index map above on random
    density, energy0 and energy1 (every cell, halos included, random before
    the fill), with only the flagged fields touched (fields[] of
    update_halo_kernel_ocl.cpp:48-62). Plus the mirror's own symmetry: after a
    full-depth fill with all faces external, reflecting the whole array
    across each wall reproduces it on every filled cell."""
    nx, ny = 13, 11
    o = oracle.OracleChunk(nx, ny, hd, 1)
    rng = np.random.default_rng(100 * hd + depth)
    names = ["density", "energy0", "energy1"]
    a0 = {n: rng.standard_normal(o.shape) for n in names}
    for n in names:
        o.put_field(n, a0[n])
    nb = [-1 if e else 3 for e in ext]
    o.update_halo((1, 0, 1), depth, chunk_neighbours=nb)
    for n in names:
        want = a0[n] if n == "energy0" else _reflected(a0[n], hd, nx, ny, depth, ext)
        assert np.array_equal(o.get_field(n), want), n
    if ext == (1, 1, 1, 1) and depth == hd:
        got = o.get_field("density")
        core = got[:, :]
        assert np.array_equal(core[:, hd - 1::-1], core[:, hd:2 * hd])
        assert np.array_equal(core[:, hd + nx:], core[:, hd + nx - 1:nx - 1:-1][:, :hd])
        assert np.array_equal(core[hd - 1::-1, :], core[hd:2 * hd, :])
        assert np.array_equal(core[hd + ny:, :], core[hd + ny - 1:ny - 1:-1, :][:hd, :])


def test_update_halo_mirror_catches_wrong_index(oracle):
    """The index map is strict enough to see a mirror that is one cell off
    (src = 2hd - j instead of 2hd - 1 - j): such a fill differs from it."""
    hd, nx, ny = 2, 9, 7
    rng = np.random.default_rng(5)
    a0 = rng.standard_normal((ny + 2 * hd, nx + 2 * hd))
    wrong = a0.copy()
    for k in range(ny + 2 * hd):
        for j in range(hd):
            wrong[k, j] = a0[k, 2 * hd - j]
    assert not np.array_equal(wrong, _reflected(a0, hd, nx, ny, hd, (1, 0, 0, 0)))


def test_operator_invariants(oracle):
    o = oracle.OracleChunk(24, 20, 1, 1)
    setup(o, seed=3)
    A = _dense_operator(o)
    assert np.allclose(A, A.T, rtol=0, atol=0)
    # A 1 = 1 through the oracle's own residual kernel (u0 = 0 -> r = -A 1)
    one = np.zeros(o.shape)
    o.interior(one)[:] = 1.0
    o.put_field("u", one)
    o.put_field("u0", np.zeros(o.shape))
    o.calc_residual()
    assert np.allclose(o.interior(o.get_field("r")), -1.0, rtol=0, atol=1e-12)
    # symmetry through the kernel
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(24 * 20), rng.standard_normal(24 * 20)
    assert abs(x @ (A @ y) - (A @ x) @ y) < 1e-10 * abs(x @ (A @ y))


def _cg_until(o, eps=1e-15, max_iters=5000):
    o.calc_residual()
    r0 = math.sqrt(o.calc_2norm(1))
    rro = o.cg_init()
    for _ in range(max_iters):
        pw = o.cg_calc_w()
        alpha = rro / pw
        rrn = o.cg_calc_ur(alpha)
        o.cg_calc_p(rrn / rro)
        rro = rrn
        if math.sqrt(abs(rrn)) < eps * r0:
            break


@pytest.mark.parametrize("prec", [1, 2, 3])
@pytest.mark.parametrize("nx,ny", [(12, 16), (10, 9)])
def test_cg_matches_dense_solve(oracle, prec, nx, ny):
    o = oracle.OracleChunk(nx, ny, 1, prec)
    setup(o, seed=4)
    A = _dense_operator(o)
    b = o.interior(o.get_field("u0")).ravel()
    want = np.linalg.solve(A, b)
    _cg_until(o)
    got = o.interior(o.get_field("u")).ravel()
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))


def test_jacobi_matches_dense_solve(oracle):
    o = oracle.OracleChunk(8, 8, 1, 1)
    setup(o, seed=2)
    A = _dense_operator(o)
    want = np.linalg.solve(A, o.interior(o.get_field("u0")).ravel())
    for _ in range(20000):
        if o.jacobi_solve() == 0.0:
            break
    got = o.interior(o.get_field("u")).ravel()
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))


@pytest.mark.parametrize("solver,extra", [
    ("tl_use_chebyshev", ["tl_ch_cg_presteps=8", "tl_ch_cg_epslim=1.0e300"]),
    ("tl_use_ppcg", ["tl_ppcg_inner_steps=6"]),
    ("tl_use_ppcg", ["tl_ppcg_inner_steps=6", "tl_ppcg_fresh_z"]),
    ("tl_use_cg", []),
])
@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
def test_deck_solvers_match_dense_solve(oracle, tmp_path, solver, extra, prec):
    """One timestep: temp = sum vol*rho*u with u the dense solution of A u = u0."""
    nx, ny = 12, 16
    states = decks.seeded_states(5, 2, False)
    deck = decks.write_deck(str(tmp_path / "d.in"), nx, ny, states, solver=solver, preconditioner=prec,
                            end_step=1, eps=1e-14, max_iters=5000, extra=extra)
    lg = oracle.run_deck(deck, tmp_path / "d.log")
    o = oracle.OracleChunk(nx, ny, 1, 1)
    setup(o, seed=5, circle=False)
    A = _dense_operator(o)
    u0 = o.interior(o.get_field("u0")).ravel()
    u = np.linalg.solve(A, u0)
    vol = (10.0 / nx) * (10.0 / ny)
    rho = o.interior(o.get_field("density")).ravel()
    temp = np.sum(vol * rho * u)
    assert abs(lg["summary"][-1][5] - temp) <= 1e-10 * abs(temp)


@pytest.mark.parametrize("nx,ny", [(16, 16), (13, 10), (7, 5)])
def test_block_jacobi_strip_systems(oracle, nx, ny):
    o = oracle.OracleChunk(nx, ny, 1, 3)
    setup(o, seed=6)
    rng = np.random.default_rng(1)
    r = np.zeros(o.shape)
    o.interior(r)[:] = rng.standard_normal((ny, nx))
    o.put_field("r", r)
    o.block_init()
    o.block_solve()
    z = o.get_field("z")
    Kx, Ky = o.get_field("Kx"), o.get_field("Ky")
    hd = 1
    for j in range(hd, hd + nx):
        for k0 in range(hd, hd + ny, 4):
            rows = [k for k in range(k0, k0 + 4) if k < hd + ny]
            for k in rows:
                b = 1.0 + Ky[k + 1, j] + Ky[k, j] + Kx[k, j + 1] + Kx[k, j]
                lhs = b * z[k, j]
                if k - 1 >= k0:
                    lhs -= Ky[k, j] * z[k - 1, j]
                if k + 1 in rows:
                    lhs -= Ky[k + 1, j] * z[k + 1, j]
                assert abs(lhs - r[k, j]) <= 1e-13 * max(1.0, abs(r[k, j]))


def test_tqli_matches_lapack(oracle):
    rng = np.random.default_rng(2)
    for n in (1, 2, 5, 30):
        d = rng.uniform(1, 10, n)
        off = rng.uniform(-1, 1, n)  # off[i] couples i-1 and i (Fortran offdiag(2..n))
        off[0] = 0.0
        T = np.diag(d) + np.diag(off[1:], 1) + np.diag(off[1:], -1)
        info, ev = oracle.tqli(d, off)
        assert info == 0
        assert np.allclose(np.sort(ev), np.linalg.eigvalsh(T), rtol=1e-12, atol=1e-12)


def test_eigenvalue_estimate_from_cg(oracle):
    """CG on diag(1..100): the Lanczos matrix of its alphas/betas spans [1, 100]."""
    lam = np.linspace(1.0, 100.0, 200)
    b = np.ones_like(lam)
    x = np.zeros_like(lam)
    r = b.copy()
    p = r.copy()
    rro = r @ r
    alphas, betas = [], []
    for _ in range(60):
        w = lam * p
        a = rro / (p @ w)
        x += a * p
        r -= a * w
        rrn = r @ r
        beta = rrn / rro
        alphas.append(a)
        betas.append(beta)
        p = r + beta * p
        rro = rrn
    info, lo, hi = oracle.calc_eigenvalues(np.array(alphas), np.array(betas))
    assert info == 0
    assert abs(lo - 1.0) < 1e-6 and abs(hi - 100.0) < 1e-6


def test_chebyshev_coefficients_closed_form(oracle):
    eigmin, eigmax, n = 0.3, 47.0, 40
    a, b, theta = oracle.calc_ch_coefs(eigmin, eigmax, n)
    th = 0.5 * (eigmax + eigmin)
    delta = 0.5 * (eigmax - eigmin)
    sigma = th / delta
    T = lambda m: math.cosh(m * math.acosh(sigma))
    assert theta == th
    rho = [T(m) / T(m + 1) for m in range(n + 1)]
    for i in range(n):
        assert abs(a[i] - rho[i + 1] * rho[i]) <= 1e-12 * abs(a[i])
        assert abs(b[i] - 2.0 * rho[i + 1] / delta) <= 1e-12 * abs(b[i])


def test_generate_chunk_geometry(oracle):
    """Cell-overlap rectangles and centre-in-circle, against a numpy painting."""
    nx, ny = 50, 40
    o = oracle.OracleChunk(nx, ny, 2, 1)
    setup(o, seed=9, n_rect=3, circle=True)
    st = decks.seeded_states(9, 3, True)
    dx, dy = 10.0 / nx, 10.0 / ny
    xv = dx * np.arange(nx + 1)
    yv = dy * np.arange(ny + 1)
    xc = 0.5 * (xv[:-1] + xv[1:])
    yc = 0.5 * (yv[:-1] + yv[1:])
    rho = np.full((ny, nx), st[0]["density"])
    for s in st[1:]:
        if s["geometry"] == "rectangle":
            x0, x1 = s["xmin"] + dx / 100, s["xmax"] - dx / 100
            y0, y1 = s["ymin"] + dy / 100, s["ymax"] - dy / 100
            mx = (xv[1:] >= x0) & (xv[:-1] < x1)
            my = (yv[1:] >= y0) & (yv[:-1] < y1)
            rho[np.outer(my, mx)] = s["density"]
        else:
            cx, cy = s["xmin"] + dx / 100, s["ymin"] + dy / 100
            m = np.hypot(xc[None, :] - cx, yc[:, None] - cy) <= s["radius"]
            rho[m] = s["density"]
    assert np.array_equal(o.interior(o.get_field("density")), rho)
