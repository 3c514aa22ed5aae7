"""tea.in deck writers: the reference's own known-answer decks and seeded
synthetic layouts.

Synthetic layout (domain [0,10]^2): state 1 is the background (density 100,
energy 1e-4); each seeded rectangle covers an area fraction f ~ U[0.05,0.20]
with aspect a ~ U[0.5,2] (w = sqrt(100 f a), h = sqrt(100 f / a)) and a lower
left corner uniform in the domain, density 0.1, energy 25; an optional circle
has area fraction U[0.01,0.05] and a centre inset by its radius. Values are
written with 17 significant digits so every deck reader sees the same doubles.
The random source is splitmix64 -> U[0,1) with 53 bits.
"""
import math

import numpy as np

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def seeded_states(seed=1, n_rect=2, circle=False, rect_density=0.1, rect_energy=25.0):
    """Background + seeded rectangles (+ circle) as deck state dicts."""
    rng = SplitMix64(seed)
    states = [dict(density=100.0, energy=1.0e-4)]
    for _ in range(n_rect):
        f = rng.uniform(0.05, 0.20)
        a = rng.uniform(0.5, 2.0)
        w = math.sqrt(100.0 * f * a)
        h = math.sqrt(100.0 * f / a)
        x0 = rng.uniform(0.0, 10.0)
        y0 = rng.uniform(0.0, 10.0)
        states.append(dict(density=rect_density, energy=rect_energy, geometry="rectangle",
                           xmin=x0, xmax=x0 + w, ymin=y0, ymax=y0 + h))
    if circle:
        f = rng.uniform(0.01, 0.05)
        r = math.sqrt(100.0 * f / math.pi)
        cx = rng.uniform(r, 10.0 - r)
        cy = rng.uniform(r, 10.0 - r)
        states.append(dict(density=rect_density, energy=rect_energy, geometry="circle",
                           xmin=cx, ymin=cy, radius=r))
    return states


KAT_STATES = [
    dict(density=100.0, energy=0.0001),
    dict(density=0.1, energy=25.0, geometry="rectangle", xmin=0.0, xmax=1.0, ymin=1.0, ymax=2.0),
    dict(density=0.1, energy=0.1, geometry="rectangle", xmin=1.0, xmax=6.0, ymin=1.0, ymax=2.0),
    dict(density=0.1, energy=0.1, geometry="rectangle", xmin=5.0, xmax=6.0, ymin=1.0, ymax=8.0),
    dict(density=0.1, energy=0.1, geometry="rectangle", xmin=5.0, xmax=10.0, ymin=7.0, ymax=8.0),
]


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


def _fmt(v):
    if isinstance(v, float):
        s = "%.17g" % v
        if "." not in s and "e" not in s and "inf" not in s and "nan" not in s:
            s += ".0"
        return s
    return str(v)


def write_deck(path, x_cells, y_cells, states, solver="tl_use_cg", preconditioner=None,
               end_step=10, dt=0.004, eps=1.0e-15, max_iters=5000, halo_depth=1,
               summary_frequency=1, coefficient="tl_coefficient_density", extra=(),
               xmax=10.0, ymax=10.0):
    """Write a tea.in deck. `preconditioner` is none/jac_diag/jac_block or None."""
    lines = ["*tea"]
    for i, st in enumerate(states, 1):
        parts = ["state %d" % i, "density=%s" % _fmt(float(st["density"])),
                 "energy=%s" % _fmt(float(st["energy"]))]
        if "geometry" in st:
            parts.append("geometry=%s" % st["geometry"])
            for k in ("xmin", "xmax", "ymin", "ymax", "radius"):
                if k in st:
                    parts.append("%s=%s" % (k, _fmt(float(st[k]))))
        # the reference reads 100 characters per line (parse.f90:154): wrap
        # long state specifications with the "\\" continuation word
        cur = parts[0]
        for w in parts[1:]:
            if len(cur) + 1 + len(w) + 2 > 100:
                lines.append(cur + " \\")
                cur = w
            else:
                cur += " " + w
        lines.append(cur)
    lines += ["x_cells=%d" % x_cells, "y_cells=%d" % y_cells,
              "xmin=0.0", "ymin=0.0", "xmax=%s" % _fmt(float(xmax)), "ymax=%s" % _fmt(float(ymax)),
              "initial_timestep=%s" % _fmt(float(dt)), "end_step=%d" % end_step,
              "tl_max_iters=%d" % max_iters, solver, "tl_eps=%s" % _fmt(float(eps)),
              coefficient, "summary_frequency=%d" % summary_frequency, "halo_depth=%d" % halo_depth]
    if preconditioner:
        lines.append("tl_preconditioner_type=%s" % preconditioner)
    lines += list(extra)
    lines.append("*endtea")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def config1_deck(path, seed=1, n=512, steps=10, preconditioner="jac_diag", solver="tl_use_cg", **kw):
    """512^2 CG + diagonal Jacobi, two seeded rectangles (the CPU-runnable case)."""
    return write_deck(path, n, n, seeded_states(seed, 2, False), solver=solver,
                      preconditioner=preconditioner, end_step=steps, **kw)


def config2_deck(path, solver="tl_use_cg", preconditioner="jac_block", seed=1, n=4096, steps=20, extra=(), **kw):
    """4096^2, two rectangles and a circle; CG + block Jacobi by default."""
    ex = list(extra)
    if solver == "tl_use_chebyshev":
        ex += ["tl_ch_cg_presteps=20", "tl_ch_cg_epslim=1.0e300"]
    if solver == "tl_use_ppcg":
        ex += ["tl_ppcg_inner_steps=10"]
    return write_deck(path, n, n, seeded_states(seed, 2, True), solver=solver,
                      preconditioner=preconditioner, end_step=steps, extra=ex, **kw)


def config3_deck(path, precision="mixed", preconditioner="jac_block", seed=1, n=4096, steps=1, eps=1e-10,
                 extra=(), **kw):
    """4096^2, the two seeded rectangles, conductivity 1/density (contrast
    1e3), CG + block Jacobi. precision: "mixed" (FP32 inner CG with FP64
    residual replacement), "fp64" (the same refined solve with FP64 inner
    vectors) or "reference" (the reference CG and its exit test)."""
    ex = list(extra)
    if precision == "mixed":
        ex.append("tl_mixed_precision")
    elif precision == "fp64":
        ex.append("tl_fp64_refined")
    elif precision != "reference":
        raise ValueError("precision must be mixed, fp64 or reference")
    return write_deck(path, n, n, seeded_states(seed, 2, False), solver="tl_use_cg",
                      preconditioner=preconditioner, end_step=steps, eps=eps,
                      coefficient="tl_coefficient_inverse_density", extra=ex, **kw)


def _splitmix_uniform(seed, n):
    """n draws of the splitmix64 -> U[0,1) stream of SplitMix64(seed), vectorised."""
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def config4_density(n, seed=1, hd=1, sigma=1.0):
    """Per-cell density of the adjoint workload on [0,10]^2 (layout of
    Chunk.get_field, halo cells zero): rho = 10^U[-1,2] * (1 + 0.5 exp(-|x - x0|^2
    / (2 sigma^2))) at the cell centres, x0 uniform in [2,8]^2, all draws from
    splitmix64(seed) (x0 first, then the cells row by row)."""
    rng = SplitMix64(seed)
    x0 = 2.0 + 6.0 * rng.uniform()
    y0 = 2.0 + 6.0 * rng.uniform()
    u = _splitmix_uniform(rng.state, n * n).reshape(n, n)
    dx = 10.0 / n
    c = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(c, c)
    bump = 1.0 + 0.5 * np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / (2.0 * sigma * sigma))
    rho = np.zeros((n + 2 * hd, n + 2 * hd))
    rho[hd:hd + n, hd:hd + n] = 10.0 ** (-1.0 + 3.0 * u) * bump
    return rho


def config5_states(seed=1, n_rect=8, rect_energy=25.0):
    """Background (density 100, energy 1e-4) and `n_rect` seeded rectangles
    (area fraction U[0.05,0.20], aspect U[0.5,2], lower-left corner uniform
    over the positions where the rectangle fits), each with its own density
    10^U[-1,2] and the given energy."""
    rng = SplitMix64(seed)
    states = [dict(density=100.0, energy=1.0e-4)]
    for _ in range(n_rect):
        f = rng.uniform(0.05, 0.20)
        a = rng.uniform(0.5, 2.0)
        w = math.sqrt(100.0 * f * a)
        h = math.sqrt(100.0 * f / a)
        x0 = rng.uniform(0.0, 10.0 - min(w, 10.0))
        y0 = rng.uniform(0.0, 10.0 - min(h, 10.0))
        rho = 10.0 ** rng.uniform(-1.0, 2.0)
        states.append(dict(density=rho, energy=rect_energy, geometry="rectangle",
                           xmin=x0, xmax=min(x0 + w, 10.0), ymin=y0, ymax=min(y0 + h, 10.0)))
    return states


def config5_deck(path, precision="reference", n=8192, steps=20, seed=1, eps=1e-10, inner=10, presteps=20, **kw):
    """The large combined run: PPCG + block Jacobi on config5_states, in the
    reference FP64 PPCG ("reference", its own exit test), or the refined PPCG
    to a true relative residual of eps with FP64 ("fp64") or FP32 ("mixed")
    inner fields."""
    ex = ["tl_ppcg_inner_steps=%d" % inner, "tl_ch_cg_presteps=%d" % presteps, "tl_ch_cg_epslim=1.0e300"]
    if precision == "mixed":
        ex.append("tl_mixed_precision")
    elif precision == "fp64":
        ex.append("tl_fp64_refined")
    elif precision != "reference":
        raise ValueError("precision must be reference, fp64 or mixed")
    return write_deck(path, n, n, config5_states(seed), solver="tl_use_ppcg", preconditioner="jac_block",
                      end_step=steps, eps=eps, extra=ex, **kw)
