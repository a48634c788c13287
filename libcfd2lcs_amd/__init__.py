"""MI355X-native libcfd2lcs FTLE path: Python binding of the C ABI.

The product is the pair of shared libraries built from ``csrc/`` for gfx950
(``lib/libcfd2lcs_dp.so`` with ``lcsdata_t = double`` and
``lib/libcfd2lcs_sp.so`` with ``lcsdata_t = float``).  This module binds their
C entry points with ctypes and mirrors the reference's API (cfd2lcs_c.c:67-142):
``cfd2lcs_init``, ``cfd2lcs_diagnostic_init``, ``cfd2lcs_update``,
``cfd2lcs_set_option``, ``cfd2lcs_set_param``, ``cfd2lcs_diagnostic_destroy``,
plus the additive result getters.  There is no CPU fallback: importing works
without a GPU, but every call that needs the device fails loudly.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

# flags (cfd2lcs_inc_double.h:27-65)
LCS_INTERNAL, LCS_INFLOW, LCS_OUTFLOW, LCS_WALL, LCS_SLIP, LCS_2D, LCS_MASK = range(7)
FTLE_FWD, FTLE_BKWD, LP_TRACER = 0, 1, 2
EULER, TRAPEZOIDAL, RK2, RK3, RK4 = 0, 1, 2, 3, 4
NEAREST_NBR, LINEAR, QUADRATIC, CUBIC, TSE, TSE_LIMIT = 0, 1, 2, 3, 5, 6
LCS_FALSE, LCS_TRUE = 0, 1
LCS_NAMELEN = 128
# additive option values ("PRECISION", "VELOCITY_STORAGE")
LCS_FP64, LCS_FP32, LCS_FP16, LCS_BF16 = 8, 4, 2, 3

EXPORTED = (
    "cfd2lcs_init_c", "cfd2lcs_diagnostic_init_c", "cfd2lcs_update_c",
    "cfd2lcs_diagnostic_destroy_c", "cfd2lcs_set_option_c", "cfd2lcs_set_param_c",
    "cfd2lcs_finalize_c", "cfd2lcs_error_c", "cfd2lcs_update_device_c",
    "cfd2lcs_get_flowmap_c", "cfd2lcs_get_composed_c", "cfd2lcs_get_ftle_c", "cfd2lcs_get_lcs_grid_c",
    "cfd2lcs_get_ftle_device_c", "cfd2lcs_get_noutput_c", "cfd2lcs_get_stats_c",
    "cfd2lcs_get_timers_c", "cfd2lcs_debug_grid_c",
    "cfd2lcs_adjoint_forward_c", "cfd2lcs_adjoint_backward_c", "cfd2lcs_adjoint_release_c",
    "cfd2lcs_adjoint_stats_c", "cfd2lcs_adjoint_forward_noalloc_c", "cfd2lcs_flush_output_c",
    "cfd2lcs_format_es_c",
    "cfd2lcs_init_", "cfd2lcs_diagnostic_init_", "cfd2lcs_update_",
    "cfd2lcs_diagnostic_destroy_", "cfd2lcs_set_option_", "cfd2lcs_set_param_",
    "cfd2lcs_finalize_",
)

_libs = {}


def lib_path(variant="dp"):
    return os.path.join(LIBDIR, f"libcfd2lcs_{variant}.so")


def load(variant="dp"):
    """Load and bind libcfd2lcs_{dp,sp}.so; raises if it has not been built."""
    if variant in _libs:
        return _libs[variant]
    path = lib_path(variant)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m libcfd2lcs_amd.build`")
    L = C.CDLL(path)
    real = C.c_double if variant == "dp" else C.c_float
    rp = C.POINTER(real)
    ip = C.POINTER(C.c_int)
    L.cfd2lcs_init_c.argtypes = [C.c_int, ip, ip, rp, rp, rp, ip]
    L.cfd2lcs_init_c.restype = None
    L.cfd2lcs_diagnostic_init_c.argtypes = [C.c_int, C.c_int, real, real, C.c_char_p]
    L.cfd2lcs_diagnostic_init_c.restype = C.c_int
    L.cfd2lcs_update_c.argtypes = [ip, rp, rp, rp, real]
    L.cfd2lcs_update_c.restype = None
    L.cfd2lcs_update_device_c.argtypes = [ip, C.c_void_p, C.c_void_p, C.c_void_p, real]
    L.cfd2lcs_update_device_c.restype = None
    L.cfd2lcs_diagnostic_destroy_c.argtypes = [C.c_int]
    L.cfd2lcs_set_option_c.argtypes = [C.c_char_p, C.c_int]
    L.cfd2lcs_set_param_c.argtypes = [C.c_char_p, real]
    L.cfd2lcs_finalize_c.argtypes = []
    L.cfd2lcs_error_c.restype = C.c_int
    for name in ("cfd2lcs_get_flowmap_c", "cfd2lcs_get_composed_c"):
        getattr(L, name).argtypes = [C.c_int, rp, rp, rp]
        getattr(L, name).restype = C.c_int
    L.cfd2lcs_get_ftle_c.argtypes = [C.c_int, rp]
    L.cfd2lcs_get_ftle_c.restype = C.c_int
    L.cfd2lcs_get_lcs_grid_c.argtypes = [C.c_int, ip, rp, rp, rp]
    L.cfd2lcs_get_lcs_grid_c.restype = C.c_int
    L.cfd2lcs_get_ftle_device_c.argtypes = [C.c_int]
    L.cfd2lcs_get_ftle_device_c.restype = C.c_void_p
    L.cfd2lcs_get_noutput_c.argtypes = [C.c_int]
    L.cfd2lcs_get_noutput_c.restype = C.c_int
    L.cfd2lcs_get_stats_c.argtypes = [C.POINTER(C.c_double)]
    L.cfd2lcs_debug_grid_c.argtypes = [C.c_void_p] * 5
    L.cfd2lcs_debug_grid_c.restype = C.c_int
    L.cfd2lcs_adjoint_forward_c.argtypes = [ip, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, rp, C.c_int,
                                            C.c_void_p]
    L.cfd2lcs_adjoint_forward_c.restype = C.c_int
    L.cfd2lcs_adjoint_forward_noalloc_c.argtypes = [ip, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, rp,
                                                    C.c_void_p]
    L.cfd2lcs_adjoint_forward_noalloc_c.restype = C.c_int
    L.cfd2lcs_adjoint_backward_c.argtypes = [C.c_void_p] * 4
    L.cfd2lcs_adjoint_backward_c.restype = C.c_int
    L.cfd2lcs_adjoint_release_c.argtypes = []
    L.cfd2lcs_adjoint_stats_c.argtypes = [C.POINTER(C.c_double)]
    L.cfd2lcs_flush_output_c.argtypes = []
    L.cfd2lcs_flush_output_c.restype = C.c_int
    L.cfd2lcs_format_es_c.argtypes = [C.POINTER(C.c_double), C.c_longlong, C.c_int, C.c_int, C.c_char_p]
    L.cfd2lcs_format_es_c.restype = C.c_int
    _libs[variant] = L
    return L


def _fortran(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))


def format_es(values, w=14, d=6, variant="dp"):
    """The Fortran ESw.d strings the output files use, one per value (host
    only; no GPU needed)."""
    L = load(variant)
    v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).ravel())
    out = C.create_string_buffer(max(1, v.size * w))
    if L.cfd2lcs_format_es_c(v.ctypes.data_as(C.POINTER(C.c_double)), v.size, w, d, out) != 0:
        raise ValueError(f"unsupported edit descriptor ES{w}.{d}")
    raw = out.raw[:v.size * w].decode()
    return [raw[i * w:(i + 1) * w] for i in range(v.size)]


class Cfd2lcs:
    """The reference's API on one GPU.  The library is process-global like the
    reference's module state; ``finalize`` resets it (cfd2lcs.f90:388-434).

    The C library writes the reference's output files (``./cfd2lcs_output``)
    at every FTLE output unless the option "WRITE_LCS" is LCS_FALSE; this
    binding sets it to LCS_FALSE in ``init`` unless ``write_output`` is true,
    since its callers read the results from memory."""

    def __init__(self, variant="dp", write_output=False):
        self.write_output = write_output
        self.variant = variant
        self.L = load(variant)
        self.dtype = np.float64 if variant == "dp" else np.float32
        self.real = C.c_double if variant == "dp" else C.c_float
        self.shape = None
        self._keep = []

    def _rp(self, a):
        return a.ctypes.data_as(C.POINTER(self.real))

    # ---- reference API ----
    def init(self, x, y, z, flag, offset=(0, 0, 0), comm=0):
        self.shape = tuple(int(v) for v in np.shape(x))
        n = (C.c_int * 3)(*self.shape)
        off = (C.c_int * 3)(*offset)
        xs, ys, zs = (_fortran(a, self.dtype) for a in (x, y, z))
        fl = _fortran(flag, np.int32)
        if not self.write_output:
            self.set_option("WRITE_LCS", LCS_FALSE)
        self.L.cfd2lcs_init_c(comm, n, off, self._rp(xs), self._rp(ys), self._rp(zs),
                              fl.ctypes.data_as(C.POINTER(C.c_int)))

    def diagnostic_init(self, lcs_type, resolution, T, h, label):
        return self.L.cfd2lcs_diagnostic_init_c(lcs_type, resolution, T, h, label.encode())

    def update(self, u, v, w, t):
        n = (C.c_int * 3)(*self.shape)
        us, vs, ws = (_fortran(a, self.dtype) for a in (u, v, w))
        self.L.cfd2lcs_update_c(n, self._rp(us), self._rp(vs), self._rp(ws), t)

    def update_device(self, du, dv, dw, t):
        """u, v, w as device pointers (ints), e.g. torch tensors' data_ptr()."""
        n = (C.c_int * 3)(*self.shape)
        self.L.cfd2lcs_update_device_c(n, C.c_void_p(du), C.c_void_p(dv), C.c_void_p(dw), t)

    def set_option(self, name, val):
        self.L.cfd2lcs_set_option_c(name.encode(), int(val))

    def set_param(self, name, val):
        self.L.cfd2lcs_set_param_c(name.encode(), val)

    def diagnostic_destroy(self, handle):
        self.L.cfd2lcs_diagnostic_destroy_c(handle)

    def finalize(self):
        self.L.cfd2lcs_finalize_c()

    def flush_output(self):
        """Wait until every output file requested so far is written."""
        return self.L.cfd2lcs_flush_output_c()

    # ---- additive ----
    @property
    def error(self):
        return self.L.cfd2lcs_error_c()

    def lcs_grid(self, handle, coords=True):
        """Extents (and node coordinates x, y, z) of a diagnostic's grid: the
        flow grid, or the refined / coarsened grid of resolution != 0."""
        n = (C.c_int * 3)()
        if self.L.cfd2lcs_get_lcs_grid_c(handle, n, None, None, None) != 0:
            raise RuntimeError(f"grid query failed for handle {handle}")
        shape = tuple(int(v) for v in n)
        if not coords:
            return shape
        xyz = [np.zeros(int(np.prod(shape)), dtype=self.dtype) for _ in range(3)]
        if self.L.cfd2lcs_get_lcs_grid_c(handle, n, *[self._rp(a) for a in xyz]) != 0:
            raise RuntimeError(f"grid query failed for handle {handle}")
        return shape, [a.reshape(shape, order="F") for a in xyz]

    def _get3(self, fn, handle):
        shape = self.lcs_grid(handle, coords=False)
        npt = int(np.prod(shape))
        out = [np.zeros(npt, dtype=self.dtype) for _ in range(3)]
        rc = fn(handle, *[self._rp(o) for o in out])
        if rc != 0:
            raise RuntimeError(f"getter failed for handle {handle}")
        return [o.reshape(shape, order="F") for o in out]

    def flowmap(self, handle):
        return self._get3(self.L.cfd2lcs_get_flowmap_c, handle)

    def composed(self, handle):
        return self._get3(self.L.cfd2lcs_get_composed_c, handle)

    def ftle(self, handle):
        shape = self.lcs_grid(handle, coords=False)
        out = np.zeros(int(np.prod(shape)), dtype=self.dtype)
        if self.L.cfd2lcs_get_ftle_c(handle, self._rp(out)) != 0:
            raise RuntimeError(f"ftle getter failed for handle {handle}")
        return out.reshape(shape, order="F")

    # ---- adjoint of the backward FTLE (device pointers) ----
    def adjoint_forward(self, du, dv, dw, nsnap, times, checkpoint, dftle):
        """du, dv, dw: device pointers to nsnap snapshots each; times: host
        sequence; dftle: device pointer to n0*n1*n2 doubles (output)."""
        n = (C.c_int * 3)(*self.shape)
        ts = np.ascontiguousarray(np.asarray(times, dtype=self.dtype))
        return self.L.cfd2lcs_adjoint_forward_c(n, int(nsnap), C.c_void_p(du), C.c_void_p(dv), C.c_void_p(dw),
                                                self._rp(ts), int(checkpoint), C.c_void_p(dftle))

    def adjoint_forward_noalloc(self, du, dv, dw, nsnap, times, dftle):
        """The FTLE of adjoint_forward without a tape, computed in the
        library's fixed device workspace (no device allocation)."""
        n = (C.c_int * 3)(*self.shape)
        ts = np.ascontiguousarray(np.asarray(times, dtype=self.dtype))
        return self.L.cfd2lcs_adjoint_forward_noalloc_c(n, int(nsnap), C.c_void_p(du), C.c_void_p(dv),
                                                        C.c_void_p(dw), self._rp(ts), C.c_void_p(dftle))

    def adjoint_backward(self, ddftle, gu, gv, gw):
        return self.L.cfd2lcs_adjoint_backward_c(C.c_void_p(ddftle), C.c_void_p(gu), C.c_void_p(gv),
                                                 C.c_void_p(gw))

    def adjoint_release(self):
        self.L.cfd2lcs_adjoint_release_c()

    def adjoint_stats(self):
        out = (C.c_double * 4)()
        self.L.cfd2lcs_adjoint_stats_c(out)
        return {"forward_s": out[0], "backward_s": out[1], "tape_bytes": out[2], "subcycles": out[3]}

    def ftle_device_ptr(self, handle):
        return self.L.cfd2lcs_get_ftle_device_c(handle)

    def noutput(self, handle):
        return self.L.cfd2lcs_get_noutput_c(handle)

    def stats(self):
        out = (C.c_double * 8)()
        self.L.cfd2lcs_get_stats_c(out)
        keys = ("nsub", "integrations_fwd", "integrations_bkwd", "t_fwd", "t_bkwd",
                "t_reconstruct", "t_ftle", "spline_out_of_range")
        return dict(zip(keys, list(out)))

    def debug_grid(self):
        N = tuple(s + 2 for s in self.shape)
        lines = np.zeros(sum(N))
        nt = int(np.prod(N))
        npt = int(np.prod(self.shape))
        flag = np.zeros(nt, dtype=np.int32)
        bn = np.zeros(3 * npt)
        ps = np.zeros(81)
        bb = np.zeros(6)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        if self.L.cfd2lcs_debug_grid_c(ptr(lines), ptr(flag), ptr(bn), ptr(ps), ptr(bb)) != 0:
            raise RuntimeError("debug_grid failed")
        l3 = [lines[:N[0]], lines[N[0]:N[0] + N[1]], lines[N[0] + N[1]:]]
        return l3, flag.reshape(N, order="F"), [bn[d * npt:(d + 1) * npt].reshape(self.shape, order="F")
                                                  for d in range(3)], ps.reshape(27, 3), bb
