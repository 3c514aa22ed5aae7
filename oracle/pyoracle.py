"""ctypes wrapper of the CPU oracle (oracle/tl_oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker or the CPU baseline. The class mirrors the
method names of tealeaf_opencl_amd.Chunk so parity tests call both alike.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libtloracle.so")

FIELDS = ["density", "energy0", "energy1", "volume", "xarea", "yarea", "u", "u0", "p", "r", "w",
          "Mi", "Kx", "Ky", "sd", "z", "cp", "bfp"]
FIELDS1D = ["cellx", "celldx", "vertexx", "vertexdx", "celly", "celldy", "vertexy", "vertexdy"]

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

_SIGS = {
    "tlo_create": (_P, [_I, _I, _I, _I]),
    "tlo_destroy": (None, [_P]),
    "tlo_field2d": (_DP, [_P, _I]),
    "tlo_field1d": (_DP, [_P, _I]),
    "tlo_pitch": (_I, [_P]),
    "tlo_field2d_len": (ctypes.c_long, [_P]),
    "tlo_set_preconditioner": (None, [_P, _I]),
    "tlo_initialise_chunk": (None, [_P, _D, _D, _D, _D]),
    "tlo_generate_chunk": (None, [_P, _I, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _IP, _I, _I, _I]),
    "tlo_set_field": (None, [_P]),
    "tlo_field_summary": (None, [_P, _DP, _DP, _DP, _DP]),
    "tlo_update_halo": (None, [_P, _IP, _IP, _I]),
    "tlo_common_init": (None, [_P, _I, _D, _DP, _DP, _IP, _I]),
    "tlo_calc_residual": (None, [_P]),
    "tlo_calc_2norm": (_D, [_P, _I]),
    "tlo_finalise": (None, [_P]),
    "tlo_cg_init": (_D, [_P]),
    "tlo_cg_calc_w": (_D, [_P]),
    "tlo_cg_calc_ur": (_D, [_P, _D]),
    "tlo_cg_calc_p": (None, [_P, _D]),
    "tlo_cheby_init": (None, [_P, _DP, _DP, _I, _D, _D, _D]),
    "tlo_cheby_iterate": (None, [_P, _I]),
    "tlo_ppcg_init": (None, [_P, _DP, _DP, _D, _I]),
    "tlo_ppcg_init_sd": (None, [_P]),
    "tlo_ppcg_inner": (None, [_P, _I, _I, _IP]),
    "tlo_jacobi_solve": (_D, [_P]),
    "tlo_block_init": (None, [_P]),
    "tlo_block_solve": (None, [_P]),
    "tlo_init_jac_diag": (None, [_P]),
    "tlo_tqli": (_I, [_DP, _DP, _I]),
    "tlo_calc_eigenvalues": (_I, [_DP, _DP, _I, _DP, _DP]),
    "tlo_calc_ch_coefs": (None, [_DP, _DP, _D, _D, _DP, _I]),
    "tlo_run_deck": (_I, [ctypes.c_char_p, ctypes.c_char_p, _I]),
    "tlo_parse_deck_values": (_I, [ctypes.c_char_p, _IP, _DP, _DP, _I]),
    "tlo_time_cg_iterations": (_D, [ctypes.c_char_p, _I, _IP]),
    "tlo_num_threads": (_I, []),
}

_lib = None


def parse_deck_values(deck, max_states=64):
    """tlo_parse_deck_values: the oracle reader's settings in the arrays of the
    library's tl_parse_deck; None when the reader rejects the deck."""
    ints = (ctypes.c_int * 23)()
    reals = np.zeros(8)
    states = np.zeros((max_states, 8))
    rc = load().tlo_parse_deck_values(str(deck).encode(), ints, reals.ctypes.data_as(_DP),
                                      states.ctypes.data_as(_DP), int(max_states))
    if rc:
        return None
    iv = list(ints)
    return iv, reals, states[:min(iv[18], max_states)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _dptr(a):
    return a.ctypes.data_as(_DP)


def _iptr(a):
    return a.ctypes.data_as(_IP)


class OracleChunk:
    """CPU chunk in the reference layout; same method names as tealeaf_opencl_amd.Chunk."""

    def __init__(self, x_max, y_max, halo_depth=1, preconditioner=1):
        self.lib = load()
        self.x_max, self.y_max, self.hd = int(x_max), int(y_max), int(halo_depth)
        self.preconditioner = int(preconditioner)
        self.h = self.lib.tlo_create(self.x_max, self.y_max, self.hd, self.preconditioner)
        if not self.h:
            raise ValueError("bad chunk geometry")
        self.pitch = self.lib.tlo_pitch(self.h)
        self.len = self.lib.tlo_field2d_len(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.tlo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.y_max + 2 * self.hd, self.x_max + 2 * self.hd)

    def _view(self, name):
        fid = FIELDS.index(name) if isinstance(name, str) else int(name)
        ptr = self.lib.tlo_field2d(self.h, fid)
        return np.ctypeslib.as_array(ptr, shape=(self.len,))

    def get_field(self, name):
        rows, cols = self.shape
        return self._view(name)[:rows * self.pitch].reshape(rows, self.pitch)[:, :cols].copy()

    def put_field(self, name, arr):
        rows, cols = self.shape
        v = self._view(name)[:rows * self.pitch].reshape(rows, self.pitch)
        v[:, :cols] = arr

    def get_field1d(self, name, n):
        fid = FIELDS1D.index(name) if isinstance(name, str) else int(name)
        ptr = self.lib.tlo_field1d(self.h, fid)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def interior(self, a):
        return a[self.hd:self.hd + self.y_max, self.hd:self.hd + self.x_max]

    def initialise_chunk(self, xmin, ymin, dx, dy):
        self.lib.tlo_initialise_chunk(self.h, xmin, ymin, dx, dy)

    def generate_chunk(self, density, energy, xmin, xmax, ymin, ymax, radius, geometry,
                       g_rect=1, g_circ=2, g_point=3):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (density, energy, xmin, xmax, ymin, ymax, radius)]
        geom = np.ascontiguousarray(geometry, dtype=np.int32)
        self.lib.tlo_generate_chunk(self.h, len(arrs[0]), *[_dptr(a) for a in arrs], _iptr(geom),
                                    g_rect, g_circ, g_point)

    def set_field(self):
        self.lib.tlo_set_field(self.h)

    def field_summary(self):
        v = [_D() for _ in range(4)]
        self.lib.tlo_field_summary(self.h, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def update_halo(self, fields, depth, chunk_neighbours=(-1, -1, -1, -1)):
        nb = np.ascontiguousarray(chunk_neighbours, dtype=np.int32)
        f = np.zeros(7, dtype=np.int32)
        f[:len(fields)] = fields
        self.lib.tlo_update_halo(self.h, _iptr(nb), _iptr(f), int(depth))

    def common_init(self, coefficient, dt, zero_boundary=(1, 1, 1, 1), reflective_boundary=False):
        rx, ry = _D(), _D()
        zb = np.ascontiguousarray(zero_boundary, dtype=np.int32)
        self.lib.tlo_common_init(self.h, int(coefficient), float(dt), ctypes.byref(rx), ctypes.byref(ry),
                                 _iptr(zb), int(bool(reflective_boundary)))
        return rx.value, ry.value

    def calc_residual(self):
        self.lib.tlo_calc_residual(self.h)

    def calc_2norm(self, norm_array):
        return self.lib.tlo_calc_2norm(self.h, int(norm_array))

    def finalise(self):
        self.lib.tlo_finalise(self.h)

    def cg_init(self):
        return self.lib.tlo_cg_init(self.h)

    def cg_calc_w(self):
        return self.lib.tlo_cg_calc_w(self.h)

    def cg_calc_ur(self, alpha):
        return self.lib.tlo_cg_calc_ur(self.h, float(alpha))

    def cg_calc_p(self, beta):
        self.lib.tlo_cg_calc_p(self.h, float(beta))

    def cheby_init(self, ch_alphas, ch_betas, n_coefs, rx, ry, theta):
        a = np.ascontiguousarray(ch_alphas, dtype=np.float64)
        b = np.ascontiguousarray(ch_betas, dtype=np.float64)
        self.lib.tlo_cheby_init(self.h, _dptr(a), _dptr(b), int(n_coefs), rx, ry, theta)

    def cheby_iterate(self, step):
        self.lib.tlo_cheby_iterate(self.h, int(step))

    def ppcg_init(self, ch_alphas, ch_betas, theta, n_inner_steps):
        a = np.ascontiguousarray(ch_alphas, dtype=np.float64)
        b = np.ascontiguousarray(ch_betas, dtype=np.float64)
        self.lib.tlo_ppcg_init(self.h, _dptr(a), _dptr(b), float(theta), int(n_inner_steps))

    def ppcg_init_sd(self):
        self.lib.tlo_ppcg_init_sd(self.h)

    def ppcg_inner(self, inner_step, bounds_extra=0, chunk_neighbours=(-1, -1, -1, -1)):
        nb = np.ascontiguousarray(chunk_neighbours, dtype=np.int32)
        self.lib.tlo_ppcg_inner(self.h, int(inner_step), int(bounds_extra), _iptr(nb))

    def jacobi_solve(self):
        return self.lib.tlo_jacobi_solve(self.h)

    def block_init(self):
        self.lib.tlo_block_init(self.h)

    def block_solve(self):
        self.lib.tlo_block_solve(self.h)

    def init_jac_diag(self):
        self.lib.tlo_init_jac_diag(self.h)


def tqli(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64).copy()
    e = np.ascontiguousarray(e, dtype=np.float64).copy()
    info = load().tlo_tqli(_dptr(d), _dptr(e), len(d))
    return info, d


def calc_eigenvalues(cg_alphas, cg_betas):
    a = np.ascontiguousarray(cg_alphas, dtype=np.float64)
    b = np.ascontiguousarray(cg_betas, dtype=np.float64)
    lo, hi = _D(), _D()
    info = load().tlo_calc_eigenvalues(_dptr(a), _dptr(b), len(a), ctypes.byref(lo), ctypes.byref(hi))
    return info, lo.value, hi.value


def calc_ch_coefs(eigmin, eigmax, n):
    a = np.zeros(n)
    b = np.zeros(n)
    th = _D()
    load().tlo_calc_ch_coefs(_dptr(a), _dptr(b), eigmin, eigmax, ctypes.byref(th), n)
    return a, b, th.value


def run_deck(deck, log_path, max_steps=0):
    from tealeaf_opencl_amd import parse_log  # log format shared with the HIP driver
    rc = load().tlo_run_deck(str(deck).encode(), str(log_path).encode(), int(max_steps))
    if rc != 0:
        raise RuntimeError("oracle run failed for %s" % deck)
    return parse_log(log_path)


def time_cg_iterations(deck, n_iters):
    done = _I(0)
    secs = load().tlo_time_cg_iterations(str(deck).encode(), int(n_iters), ctypes.byref(done))
    return secs, done.value


def num_threads():
    return load().tlo_num_threads()
