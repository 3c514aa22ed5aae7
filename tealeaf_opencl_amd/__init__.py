"""MI355X (gfx950) HIP implementation of TeaLeaf's implicit diffusion solve.

Python mirror of the reference's kernel entry points (tea_leaf_kernel_ocl.cpp
and friends) over the C-ABI of libtealeaf_mi355x.so. The HIP library is the
only compute path: importing works without it, but every call that needs it
raises TealeafLibraryError when it has not been built.

    from tealeaf_opencl_amd import Chunk, run_deck
    c = Chunk(512, 512, halo_depth=1, preconditioner=PREC_JAC_DIAG)
    c.initialise_chunk(0.0, 0.0, 10/512, 10/512)
"""
import ctypes
import os

import numpy as np

from .build import lib_path

__all__ = [
    "Chunk", "run_deck", "bench_deck", "parse_log", "load_library", "TealeafError",
    "TealeafLibraryError", "PREC_NONE", "PREC_JAC_DIAG", "PREC_JAC_BLOCK",
    "SOLVER_JACOBI", "SOLVER_CG", "SOLVER_CHEBY", "SOLVER_PPCG",
    "COEF_CONDUCTIVITY", "COEF_RECIP_CONDUCTIVITY", "MODE_DROPIN", "MODE_NATIVE", "FIELDS", "FIELDS1D",
    "PRECISION_FP64", "PRECISION_MIXED",
]

PREC_NONE, PREC_JAC_DIAG, PREC_JAC_BLOCK = 1, 2, 3
SOLVER_JACOBI, SOLVER_CG, SOLVER_CHEBY, SOLVER_PPCG = 1, 2, 3, 4
COEF_CONDUCTIVITY, COEF_RECIP_CONDUCTIVITY = 1, 2
MODE_DROPIN, MODE_NATIVE = 0, 1
TUNE_PPCG_STEPS_PER_LAUNCH, TUNE_CHEBY_STEPS_PER_LAUNCH = 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
G_RECT, G_CIRC, G_POINT = 1, 2, 3
# TL_NOALLOC_MAX_STATES: the most states tl_generate_chunk_noalloc stages
NOALLOC_MAX_STATES = 512

FIELDS = ["density", "energy0", "energy1", "volume", "xarea", "yarea", "u", "u0", "p", "r", "w",
          "Mi", "Kx", "Ky", "sd", "z", "cp", "bfp"]
FIELDS1D = ["cellx", "celldx", "vertexx", "vertexdx", "celly", "celldy", "vertexy", "vertexdy"]


class TealeafError(RuntimeError):
    """A call into the library returned an error status."""


class TealeafLibraryError(TealeafError):
    """The HIP library is not built or cannot be loaded."""


class BenchResult(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("iterations", ctypes.c_long),
                ("inner_iterations", ctypes.c_long), ("k1_ms", ctypes.c_double),
                ("k2_ms", ctypes.c_double), ("k_samples", ctypes.c_long),
                ("final_summary", ctypes.c_double * 5)]


class RefineResult(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("initial_residual", ctypes.c_double),
                ("rel_residual", ctypes.c_double), ("k1_ms", ctypes.c_double), ("k2_ms", ctypes.c_double),
                ("iterations", ctypes.c_int), ("updates", ctypes.c_int), ("converged", ctypes.c_int),
                ("stalled", ctypes.c_int), ("k_samples", ctypes.c_int), ("inner_iterations", ctypes.c_int)]


class AdjointResult(ctypes.Structure):
    _fields_ = [("loss", ctypes.c_double), ("forward_s", ctypes.c_double), ("adjoint_s", ctypes.c_double),
                ("work_bytes", ctypes.c_double), ("chunk_bytes", ctypes.c_double),
                ("forward_iterations", ctypes.c_long), ("recompute_iterations", ctypes.c_long),
                ("adjoint_iterations", ctypes.c_long), ("checkpoints", ctypes.c_int),
                ("stalled_solves", ctypes.c_int)]


_LIB = None
_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

_SIGS = {
    "tl_last_error": (ctypes.c_char_p, []),
    "tl_device_count": (_I, [_IP]),
    "tl_chunk_create": (_I, [_I, _I, _I, _I, _I, ctypes.POINTER(_P)]),
    "tl_chunk_destroy": (_I, [_P]),
    "tl_chunk_set_default": (_I, [_P]),
    "tl_chunk_default": (_P, []),
    "tl_field_get": (_I, [_P, _I, _DP]),
    "tl_field_set": (_I, [_P, _I, _DP]),
    "tl_field1d_get": (_I, [_P, _I, _DP, _I]),
    "tl_initialise_chunk": (_I, [_P, _D, _D, _D, _D]),
    "tl_generate_chunk": (_I, [_P, _I, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _IP, _I, _I, _I]),
    "tl_generate_chunk_noalloc": (_I, [_P, _I, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _IP, _I, _I, _I]),
    "tl_set_field": (_I, [_P]),
    "tl_field_summary": (_I, [_P, _DP, _DP, _DP, _DP]),
    "tl_update_halo": (_I, [_P, _IP, _IP, _I]),
    "tl_common_init": (_I, [_P, _I, _D, _DP, _DP, _IP, _I]),
    "tl_calc_residual": (_I, [_P]),
    "tl_calc_2norm": (_I, [_P, _I, _DP]),
    "tl_finalise": (_I, [_P]),
    "tl_cg_init": (_I, [_P, _DP]),
    "tl_cg_calc_w": (_I, [_P, _DP]),
    "tl_cg_calc_ur": (_I, [_P, _D, _DP]),
    "tl_cg_calc_p": (_I, [_P, _D]),
    "tl_cheby_init": (_I, [_P, _DP, _DP, _I, _D, _D, _D]),
    "tl_cheby_iterate": (_I, [_P, _I]),
    "tl_ppcg_init": (_I, [_P, _DP, _DP, _D, _I]),
    "tl_ppcg_init_sd": (_I, [_P]),
    "tl_ppcg_inner": (_I, [_P, _I, _I, _IP]),
    "tl_jacobi_solve": (_I, [_P, _DP]),
    "tl_block_init": (_I, [_P]),
    "tl_block_solve": (_I, [_P]),
    "tl_init_jac_diag": (_I, [_P]),
    "tl_synchronize": (_I, [_P]),
    "tl_print_profiling": (_I, [_P]),
    "tl_run_deck": (_I, [ctypes.c_char_p, ctypes.c_char_p, _I, _I]),
    "tl_run_deck_out": (_I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _I, _I]),
    "tl_bench_deck": (_I, [ctypes.c_char_p, _I, _I, _I, ctypes.POINTER(BenchResult)]),
    "tl_solve_refined": (_I, [_P, _I, _D, _I, _D, ctypes.POINTER(RefineResult)]),
    "tl_forward_loss": (_I, [_P, _I, _D, _I, _D, _I, _DP, _DP]),
    "tl_stream_bandwidth": (_I, [_P, _I, _I, _DP]),
    "tl_check_division": (_I, [_DP, _DP, _DP, _DP, ctypes.c_long]),
    "tl_parse_deck": (_I, [ctypes.c_char_p, _IP, _DP, _DP, _I]),
    "tl_calc_eigenvalues": (_I, [_DP, _DP, _I, _DP, _DP]),
    "tl_format_fortran": (_I, [_I, _D, _I, _I, ctypes.c_char_p, _I]),
    "tl_calc_ch_coefs": (_I, [_D, _D, _I, _DP, _DP, _DP]),
    "tl_set_tuning": (_I, [_P, _I, _I]),
    "tl_ppcg_inner_fused": (_I, [_P, _I, _I]),
    "tl_set_progress": (_I, [_P, _P, _P]),
    "tl_solve_refined_ppcg": (_I, [_P, _I, _D, _I, _D, _I, _I, ctypes.POINTER(RefineResult)]),
    "tl_adjoint_density": (_I, [_P, _I, _D, _I, _D, _I, _I, _DP, _DP, ctypes.POINTER(AdjointResult)]),
}


def load_library(path=None):
    """Load libtealeaf_mi355x.so (built in-tree); raise if it is missing."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    path = path or lib_path()
    if not os.path.exists(path):
        raise TealeafLibraryError(
            "libtealeaf_mi355x.so is not built (%s); run python -m tealeaf_opencl_amd.build" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _check(rc, what):
    if rc != 0:
        msg = _LIB.tl_last_error().decode(errors="replace")
        raise TealeafError("%s failed (status %d): %s" % (what, rc, msg))


def _dptr(a):
    return a.ctypes.data_as(_DP)


def _iptr(a):
    return a.ctypes.data_as(_IP)


def device_count():
    lib = load_library()
    n = _I(0)
    lib.tl_device_count(ctypes.byref(n))
    return n.value


class Chunk:
    """One TeaLeaf chunk on the GPU (types.hpp:42-367 CloverChunk).

    Methods carry the reference entry points' names and argument meaning
    (tea_leaf_kernel_ocl.cpp etc.); scalar-returning ones block until the
    value is on the host.
    """

    def __init__(self, x_max, y_max, halo_depth=1, preconditioner=PREC_NONE, profiler_on=False):
        self.lib = load_library()
        self.x_max, self.y_max, self.hd = int(x_max), int(y_max), int(halo_depth)
        self.preconditioner = int(preconditioner)
        h = _P()
        _check(self.lib.tl_chunk_create(self.x_max, self.y_max, self.hd, self.preconditioner,
                                        int(bool(profiler_on)), ctypes.byref(h)), "tl_chunk_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.tl_chunk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        """(rows, columns) of a 2-D field in the reference layout."""
        return (self.y_max + 2 * self.hd, self.x_max + 2 * self.hd)

    # ---- field transfer -------------------------------------------------
    def get_field(self, name):
        fid = FIELDS.index(name) if isinstance(name, str) else int(name)
        out = np.zeros(self.shape, dtype=np.float64)
        _check(self.lib.tl_field_get(self.h, fid, _dptr(out)), "tl_field_get")
        return out

    def put_field(self, name, arr):
        fid = FIELDS.index(name) if isinstance(name, str) else int(name)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError("field shape %s, expected %s" % (a.shape, self.shape))
        _check(self.lib.tl_field_set(self.h, fid, _dptr(a)), "tl_field_set")

    def get_field1d(self, name, n):
        fid = FIELDS1D.index(name) if isinstance(name, str) else int(name)
        out = np.zeros(n, dtype=np.float64)
        _check(self.lib.tl_field1d_get(self.h, fid, _dptr(out), int(n)), "tl_field1d_get")
        return out

    def interior(self, a):
        return a[self.hd:self.hd + self.y_max, self.hd:self.hd + self.x_max]

    # ---- reference entry points ----------------------------------------
    def initialise_chunk(self, xmin, ymin, dx, dy):
        _check(self.lib.tl_initialise_chunk(self.h, xmin, ymin, dx, dy), "initialise_chunk")

    def generate_chunk(self, density, energy, xmin, xmax, ymin, ymax, radius, geometry,
                       g_rect=G_RECT, g_circ=G_CIRC, g_point=G_POINT, noalloc=False):
        """tl_generate_chunk, or with noalloc tl_generate_chunk_noalloc (the
        state arrays staged without a device allocation)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (density, energy, xmin, xmax, ymin, ymax, radius)]
        geom = np.ascontiguousarray(geometry, dtype=np.int32)
        n = len(arrs[0])
        fn = self.lib.tl_generate_chunk_noalloc if noalloc else self.lib.tl_generate_chunk
        _check(fn(self.h, n, *[_dptr(a) for a in arrs], _iptr(geom), g_rect, g_circ, g_point), "generate_chunk")

    def set_field(self):
        _check(self.lib.tl_set_field(self.h), "set_field")

    def field_summary(self):
        v = [_D() for _ in range(4)]
        _check(self.lib.tl_field_summary(self.h, *[ctypes.byref(x) for x in v]), "field_summary")
        return tuple(x.value for x in v)

    def update_halo(self, fields, depth, chunk_neighbours=(-1, -1, -1, -1)):
        nb = np.ascontiguousarray(chunk_neighbours, dtype=np.int32)
        f = np.zeros(7, dtype=np.int32)
        f[:len(fields)] = fields
        _check(self.lib.tl_update_halo(self.h, _iptr(nb), _iptr(f), int(depth)), "update_halo")

    def common_init(self, coefficient, dt, zero_boundary=(1, 1, 1, 1), reflective_boundary=False):
        rx, ry = _D(), _D()
        zb = np.ascontiguousarray(zero_boundary, dtype=np.int32)
        _check(self.lib.tl_common_init(self.h, int(coefficient), float(dt), ctypes.byref(rx),
                                       ctypes.byref(ry), _iptr(zb), int(bool(reflective_boundary))),
               "common_init")
        return rx.value, ry.value

    def calc_residual(self):
        _check(self.lib.tl_calc_residual(self.h), "calc_residual")

    def calc_2norm(self, norm_array):
        v = _D()
        _check(self.lib.tl_calc_2norm(self.h, int(norm_array), ctypes.byref(v)), "calc_2norm")
        return v.value

    def finalise(self):
        _check(self.lib.tl_finalise(self.h), "finalise")

    def cg_init(self):
        v = _D()
        _check(self.lib.tl_cg_init(self.h, ctypes.byref(v)), "cg_init")
        return v.value

    def cg_calc_w(self):
        v = _D()
        _check(self.lib.tl_cg_calc_w(self.h, ctypes.byref(v)), "cg_calc_w")
        return v.value

    def cg_calc_ur(self, alpha):
        v = _D()
        _check(self.lib.tl_cg_calc_ur(self.h, float(alpha), ctypes.byref(v)), "cg_calc_ur")
        return v.value

    def cg_calc_p(self, beta):
        _check(self.lib.tl_cg_calc_p(self.h, float(beta)), "cg_calc_p")

    def cheby_init(self, ch_alphas, ch_betas, n_coefs, rx, ry, theta):
        a = np.ascontiguousarray(ch_alphas, dtype=np.float64)
        b = np.ascontiguousarray(ch_betas, dtype=np.float64)
        _check(self.lib.tl_cheby_init(self.h, _dptr(a), _dptr(b), int(n_coefs), rx, ry, theta), "cheby_init")

    def cheby_iterate(self, cheby_calc_step):
        _check(self.lib.tl_cheby_iterate(self.h, int(cheby_calc_step)), "cheby_iterate")

    def ppcg_init(self, ch_alphas, ch_betas, theta, n_inner_steps):
        a = np.ascontiguousarray(ch_alphas, dtype=np.float64)
        b = np.ascontiguousarray(ch_betas, dtype=np.float64)
        _check(self.lib.tl_ppcg_init(self.h, _dptr(a), _dptr(b), float(theta), int(n_inner_steps)), "ppcg_init")

    def ppcg_init_sd(self):
        _check(self.lib.tl_ppcg_init_sd(self.h), "ppcg_init_sd")

    def ppcg_inner(self, inner_step, bounds_extra=0, chunk_neighbours=(-1, -1, -1, -1)):
        nb = np.ascontiguousarray(chunk_neighbours, dtype=np.int32)
        _check(self.lib.tl_ppcg_inner(self.h, int(inner_step), int(bounds_extra), _iptr(nb)), "ppcg_inner")

    def ppcg_inner_fused(self, first_step, nsteps):
        """Inner steps first_step .. first_step+nsteps-1 with the fused kernels
        of the native driver (tl_ppcg_inner_fused)."""
        _check(self.lib.tl_ppcg_inner_fused(self.h, int(first_step), int(nsteps)), "ppcg_inner_fused")

    def jacobi_solve(self):
        v = _D()
        _check(self.lib.tl_jacobi_solve(self.h, ctypes.byref(v)), "jacobi_solve")
        return v.value

    def block_init(self):
        _check(self.lib.tl_block_init(self.h), "block_init")

    def block_solve(self):
        _check(self.lib.tl_block_solve(self.h), "block_solve")

    def init_jac_diag(self):
        _check(self.lib.tl_init_jac_diag(self.h), "init_jac_diag")

    def synchronize(self):
        _check(self.lib.tl_synchronize(self.h), "synchronize")

    _PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int)

    def set_progress(self, fn):
        """Call fn(phase, step, iterations) after every solve of the long
        calls (forward_loss, adjoint_density); None switches it off."""
        if fn is None:
            self._progress_cb = None
            _check(self.lib.tl_set_progress(self.h, None, None), "set_progress")
            return
        self._progress_cb = self._PROGRESS(lambda user, phase, step, its: fn(phase.decode(), step, its))
        _check(self.lib.tl_set_progress(self.h, ctypes.cast(self._progress_cb, ctypes.c_void_p), None),
               "set_progress")

    def stream_bandwidth(self, kind=1, reps=20):
        """Attainable streaming bytes/s on this chunk's layout (kind 0 copy,
        kind 1 the 4-read/3-write mix of the CG update kernel)."""
        v = _D()
        _check(self.lib.tl_stream_bandwidth(self.h, int(kind), int(reps), ctypes.byref(v)), "stream_bandwidth")
        return v.value

    def forward_loss(self, coefficient, dt, nsteps, u_obs, tol=1e-12, max_iters=20000):
        """L = sum over the interior of (u - u_obs)^2 after `nsteps` timesteps
        from the current energy1 (tl_forward_loss)."""
        uo = np.ascontiguousarray(u_obs, dtype=np.float64)
        if uo.shape != self.shape:
            raise ValueError("u_obs must have shape %s" % (self.shape,))
        v = _D()
        _check(self.lib.tl_forward_loss(self.h, int(coefficient), float(dt), int(nsteps), float(tol),
                                        int(max_iters), _dptr(uo), ctypes.byref(v)), "forward_loss")
        return v.value

    def adjoint_density(self, coefficient, dt, nsteps, u_obs, tol=1e-12, max_iters=20000, checkpoint_every=5):
        """(L, dL/d density, info) by the adjoint of `nsteps` timesteps
        (tl_adjoint_density); the gradient has the layout of get_field."""
        uo = np.ascontiguousarray(u_obs, dtype=np.float64)
        if uo.shape != self.shape:
            raise ValueError("u_obs must have shape %s" % (self.shape,))
        grad = np.zeros(self.shape)
        r = AdjointResult()
        _check(self.lib.tl_adjoint_density(self.h, int(coefficient), float(dt), int(nsteps), float(tol),
                                           int(max_iters), int(checkpoint_every), _dptr(uo), _dptr(grad),
                                           ctypes.byref(r)), "adjoint_density")
        info = {k: getattr(r, k) for k, _ in AdjointResult._fields_}
        return r.loss, grad, info

    def solve_refined(self, precision=PRECISION_MIXED, tol=1e-10, max_iters=10000, delta=None):
        """Solve A u = u0 after common_init to a true relative residual `tol`
        (FP32 or FP64 inner CG with FP64 residual replacement; see
        tl_solve_refined in include/tealeaf_mi355x.h)."""
        if delta is None:
            delta = 0.1 if precision == PRECISION_MIXED else 0.0
        r = RefineResult()
        _check(self.lib.tl_solve_refined(self.h, int(precision), float(tol), int(max_iters), float(delta),
                                         ctypes.byref(r)), "solve_refined")
        return _refine_dict(r)

    def solve_refined_ppcg(self, precision=PRECISION_MIXED, tol=1e-10, max_iters=10000, delta=None, presteps=20,
                           inner_steps=10):
        """The refined solve with PPCG outer steps after `presteps` CG
        iterations (tl_solve_refined_ppcg)."""
        if delta is None:
            delta = 0.1 if precision == PRECISION_MIXED else 0.0
        r = RefineResult()
        _check(self.lib.tl_solve_refined_ppcg(self.h, int(precision), float(tol), int(max_iters), float(delta),
                                              int(presteps), int(inner_steps), ctypes.byref(r)),
               "solve_refined_ppcg")
        return _refine_dict(r)


def _refine_dict(r):
    return {"seconds": r.seconds, "initial_residual": r.initial_residual, "rel_residual": r.rel_residual,
            "iterations": r.iterations, "updates": r.updates, "converged": bool(r.converged),
            "stalled": bool(r.stalled), "k1_ms": r.k1_ms, "k2_ms": r.k2_ms, "k_samples": r.k_samples,
            "inner_iterations": r.inner_iterations}


def parse_log(path):
    """Parse the driver log written by tl_run_deck (and by the oracle)."""
    out = {"summary": [], "solve": [], "switch": [], "estitc": [], "resid": [], "qa": None, "initres": [],
           "refine": []}
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "SUMMARY":
                out["summary"].append([int(t[1])] + [float(x) for x in t[2:7]])
            elif t[0] == "SOLVE":
                out["solve"].append([int(t[1]), int(t[2]), float(t[3]), float(t[4])] + [int(x) for x in t[5:10]])
            elif t[0] == "SWITCH":
                out["switch"].append([int(t[1]), int(t[2])] + [float(x) for x in t[3:8]])
            elif t[0] == "ESTITC":
                out["estitc"].append([int(t[1]), int(t[2])])
            elif t[0] == "RESID":
                out["resid"].append([int(t[1]), int(t[2]), float(t[3])])
            elif t[0] == "INITRES":
                out["initres"].append([int(t[1]), float(t[2])])
            elif t[0] == "REFINE":
                out["refine"].append([int(t[1]), int(t[2]), int(t[3]), int(t[4]), float(t[5]), float(t[6])])
            elif t[0] == "QA":
                out["qa"] = (int(t[1]), float(t[2]), t[3])
    return out


def run_deck(deck, log_path, mode=MODE_NATIVE, max_steps=0, tea_out=None):
    """Run a whole tea.in deck on the GPU; returns the parsed log. With
    `tea_out`, the reference's tea.out lines are written there too."""
    lib = load_library()
    _check(lib.tl_run_deck_out(str(deck).encode(), str(log_path).encode(),
                               str(tea_out).encode() if tea_out else None, int(mode), int(max_steps)),
           "tl_run_deck")
    return parse_log(log_path)


def parse_deck(deck, max_states=64):
    """tl_parse_deck: the deck's settings as the library reads them, without a
    GPU: (ints[23], reals[8], states[nstates, 8]) in the order of
    include/tealeaf_mi355x.h."""
    import numpy as np
    lib = load_library()
    ints = (_I * 23)()
    reals = np.zeros(8)
    states = np.zeros((max_states, 8))
    _check(lib.tl_parse_deck(str(deck).encode(), ints, _dptr(reals), _dptr(states), int(max_states)),
           "tl_parse_deck")
    iv = list(ints)
    return iv, reals, states[:min(iv[18], max_states)]


def format_fortran(kind, v, w=0, d=0):
    """tl_format_fortran: v as tea.out writes it (kind 0 Ew.d, 1 1PEw.d, 2
    Fw.d, 3 list-directed)."""
    lib = load_library()
    buf = ctypes.create_string_buffer(128)
    _check(lib.tl_format_fortran(int(kind), float(v), int(w), int(d), buf, 128), "tl_format_fortran")
    return buf.value.decode()


def calc_eigenvalues(cg_alphas, cg_betas):
    """tl_calc_eigenvalues: (eigmin, eigmax) of the Lanczos matrix of the CG
    steps' alphas and betas (tea_leaf_cheby.f90:96-143)."""
    import numpy as np
    lib = load_library()
    a = np.ascontiguousarray(cg_alphas, dtype=np.float64)
    b = np.ascontiguousarray(cg_betas, dtype=np.float64)
    lo, hi = _D(), _D()
    _check(lib.tl_calc_eigenvalues(_dptr(a), _dptr(b), a.size, ctypes.byref(lo), ctypes.byref(hi)),
           "tl_calc_eigenvalues")
    return lo.value, hi.value


def calc_ch_coefs(eigmin, eigmax, n):
    """tl_calc_ch_coefs: (alphas, betas, theta) of n Chebyshev steps
    (tea_leaf_cheby.f90:145-172)."""
    import numpy as np
    lib = load_library()
    a, b, th = np.zeros(n), np.zeros(n), _D()
    _check(lib.tl_calc_ch_coefs(float(eigmin), float(eigmax), int(n), _dptr(a), _dptr(b), ctypes.byref(th)),
           "tl_calc_ch_coefs")
    return a, b, th.value


def check_division(num, den):
    """tl_check_division: (IEEE quotient, the kernels' reciprocal-Newton
    quotient) of two float64 arrays, both computed on the GPU."""
    import numpy as np
    lib = load_library()
    num = np.ascontiguousarray(num, dtype=np.float64)
    den = np.ascontiguousarray(den, dtype=np.float64)
    if num.shape != den.shape:
        raise ValueError("num and den differ in shape")
    quot, fast = np.empty_like(num), np.empty_like(num)
    _check(lib.tl_check_division(_dptr(num), _dptr(den), _dptr(quot), _dptr(fast), num.size), "tl_check_division")
    return quot, fast


def set_tuning(key, value, chunk=None):
    """tl_set_tuning: on `chunk`, or (chunk None) the default of chunks
    created afterwards, including those of run_deck / bench_deck."""
    lib = load_library()
    _check(lib.tl_set_tuning(chunk.h if chunk is not None else None, int(key), int(value)), "tl_set_tuning")


def bench_deck(deck, mode=MODE_NATIVE, warmup=1, steps=1):
    """Timed run of a deck (see tl_bench_deck in include/tealeaf_mi355x.h)."""
    lib = load_library()
    r = BenchResult()
    _check(lib.tl_bench_deck(str(deck).encode(), int(mode), int(warmup), int(steps), ctypes.byref(r)),
           "tl_bench_deck")
    return {"seconds": r.seconds, "iterations": r.iterations, "inner_iterations": r.inner_iterations,
            "k1_ms": r.k1_ms, "k2_ms": r.k2_ms, "k_samples": r.k_samples,
            "final_summary": list(r.final_summary)}
