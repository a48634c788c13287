"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker for the GPU library and is
never imported by the product package.
"""
import ctypes as C
import os

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = os.path.join(_ROOT, "oracle", "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(_LIB)
        L.oc_init.restype = C.c_void_p
        L.oc_init.argtypes = [_ip, _dp, _dp, _dp, _ip]
        L.oc_free.argtypes = [C.c_void_p]
        L.oc_diagnostic_init.restype = C.c_int
        L.oc_diagnostic_init.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_char_p]
        L.oc_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.oc_set_param.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.oc_update.argtypes = [C.c_void_p, _ip, _dp, _dp, _dp, C.c_double]
        L.oc_error.restype = C.c_int
        L.oc_error.argtypes = [C.c_void_p]
        for name in ("oc_get_flowmap", "oc_get_composed"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.oc_get_ftle.restype = C.c_int
        L.oc_get_ftle.argtypes = [C.c_void_p, C.c_int, _dp]
        L.oc_get_lcs_grid.restype = C.c_int
        L.oc_get_lcs_grid.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oc_get_noutput.restype = C.c_int
        L.oc_get_noutput.argtypes = [C.c_void_p, C.c_int]
        L.oc_get_nsub.restype = C.c_int
        L.oc_get_nsub.argtypes = [C.c_void_p]
        L.oc_get_integrations.restype = C.c_longlong
        L.oc_get_integrations.argtypes = [C.c_void_p, C.c_int]
        L.oc_get_ghost_grid.restype = C.c_int
        L.oc_get_ghost_grid.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        L.oc_get_periodic_shift.restype = C.c_int
        L.oc_get_periodic_shift.argtypes = [C.c_void_p, _dp, _dp]
        L.oc_bspline_knots.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.oc_bspline_fit3.restype = C.c_int
        L.oc_bspline_fit3.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int,
                                      C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.oc_bspline_eval3.restype = C.c_double
        L.oc_bspline_eval3.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp,
                                       C.POINTER(C.c_int)]
        L.oc_eig_sym3_max.restype = C.c_double
        L.oc_eig_sym3_max.argtypes = [_dp]
        L.oc_cauchy_green.argtypes = [_dp, _dp]
        L.oc_velocity_invariants.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _dp]
        L.oc_track_point.restype = C.c_int
        L.oc_track_point.argtypes = [C.c_void_p, _dp, _ip, C.c_int]
        L.oc_lp_bc_point.restype = C.c_int
        L.oc_lp_bc_point.argtypes = [C.c_void_p, _dp, _ip]
        L.oc_interp_points.restype = C.c_int
        L.oc_interp_points.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_long, _dp, _ip, _dp]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


class Oracle:
    """Mirror of the reference's C API (cfd2lcs_c.c:67-142) on the oracle."""

    def __init__(self, x, y, z, flag):
        self.shape = x.shape
        self.n = np.array(x.shape, dtype=np.int32)
        fl = np.ascontiguousarray(np.asarray(flag, dtype=np.int32).ravel(order="F"))
        self.h = lib().oc_init(self.n, _f(x), _f(y), _f(z), fl)

    def close(self):
        if self.h:
            lib().oc_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def diagnostic_init(self, lcs_type, resolution, T, h, label):
        return lib().oc_diagnostic_init(self.h, lcs_type, resolution, T, h, label.encode())

    def set_option(self, name, val):
        lib().oc_set_option(self.h, name.encode(), int(val))

    def set_param(self, name, val):
        lib().oc_set_param(self.h, name.encode(), float(val))

    def update(self, u, v, w, t):
        lib().oc_update(self.h, self.n, _f(u), _f(v), _f(w), float(t))

    @property
    def error(self):
        return lib().oc_error(self.h)

    def lcs_grid(self, handle, aux_set=0):
        """extents and interior coordinates (x, y, z) of a diagnostic's grid
        (aux_set 1..6: the AUX_GRID sets X0, Y0, Z0, X1, Y1, Z1)"""
        n = np.zeros(3, dtype=np.int32)
        assert lib().oc_get_lcs_grid(self.h, handle, aux_set, n, None, None, None) == 0
        shape = tuple(int(a) for a in n)
        xyz = [np.zeros(int(np.prod(shape))) for _ in range(3)]
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        assert lib().oc_get_lcs_grid(self.h, handle, aux_set, n, *[ptr(a) for a in xyz]) == 0
        return shape, [a.reshape(shape, order="F") for a in xyz]

    def _shape_of(self, handle):
        return self.lcs_grid(handle)[0]

    def _get3(self, fn, handle):
        shape = self._shape_of(handle)
        out = [np.zeros(int(np.prod(shape))) for _ in range(3)]
        rc = fn(self.h, handle, *out)
        assert rc == 0
        return [o.reshape(shape, order="F") for o in out]

    def flowmap(self, handle):
        return self._get3(lib().oc_get_flowmap, handle)

    def composed(self, handle):
        return self._get3(lib().oc_get_composed, handle)

    def ftle(self, handle):
        shape = self._shape_of(handle)
        out = np.zeros(int(np.prod(shape)))
        assert lib().oc_get_ftle(self.h, handle, out) == 0
        return out.reshape(shape, order="F")

    def noutput(self, handle):
        return lib().oc_get_noutput(self.h, handle)

    def invariants(self, u=None, v=None, w=None):
        """compute_invariants of the velocity (u, v, w) at T0 on this grid
        (None: the zero field); returns Q, C, D, H, L2."""
        npt = int(np.prod(self.shape))
        arrs = [None if a is None else _f(a) for a in (u, v, w)]
        ptr = [None if a is None else a.ctypes.data_as(C.c_void_p) for a in arrs]
        out = np.zeros(5 * npt)
        lib().oc_velocity_invariants(self.h, *ptr, out)
        return [out[q * npt:(q + 1) * npt].reshape(self.shape, order="F") for q in range(5)]

    @property
    def nsub(self):
        return lib().oc_get_nsub(self.h)

    def integrations(self, direction):
        return lib().oc_get_integrations(self.h, direction)

    def ghost_grid(self):
        N = tuple(int(a) + 2 for a in self.shape)
        nt = int(np.prod(N))
        arrs = [np.zeros(nt) for _ in range(3)]
        fl = np.zeros(nt, dtype=np.int32)
        bn = [np.zeros(nt) for _ in range(3)]
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        rc = lib().oc_get_ghost_grid(self.h, *[ptr(a) for a in arrs], ptr(fl), *[ptr(a) for a in bn])
        assert rc == 0
        rs = lambda a: a.reshape(N, order="F")
        return [rs(a) for a in arrs], rs(fl), [rs(a) for a in bn]

    def track(self, q, node, recursive=True):
        """track_lp2node of point q from 1-based node (i, j, k) on the flow
        grid; returns (node, passes)"""
        n3 = np.ascontiguousarray(node, dtype=np.int32).copy()
        r = lib().oc_track_point(self.h, np.ascontiguousarray(q, dtype=np.float64), n3, int(recursive))
        return tuple(int(a) for a in n3), r

    def lp_bc(self, q, node):
        """set_lp_bc of point q tracked to node; returns (q, node, stick)"""
        qq = np.ascontiguousarray(q, dtype=np.float64).copy()
        n3 = np.ascontiguousarray(node, dtype=np.int32).copy()
        r = lib().oc_lp_bc_point(self.h, qq, n3)
        assert r >= 0
        return qq, tuple(int(a) for a in n3), bool(r)

    def interp(self, interp, u, v, w, q, nodes):
        """interp_s2u_r1 of the interior field (u, v, w) at points q (npt, 3)
        from nodes (npt, 3), after the velocity ghost exchange and BC"""
        q = np.ascontiguousarray(np.atleast_2d(q), dtype=np.float64)
        nd = np.ascontiguousarray(np.atleast_2d(nodes), dtype=np.int32)
        out = np.zeros(q.shape)
        rc = lib().oc_interp_points(self.h, int(interp), _f(u), _f(v), _f(w), q.shape[0], q.ravel(),
                                    nd.ravel(), out.ravel())
        assert rc == 0
        return out

    def periodic_shift(self):
        ps = np.zeros(81)
        bb = np.zeros(6)
        lib().oc_get_periodic_shift(self.h, ps, bb)
        return ps.reshape(27, 3), bb


def bspline_knots(x, k):
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.zeros(len(x) + k)
    lib().oc_bspline_knots(x, len(x), k, t)
    return t


def bspline_fit3(x, y, z, f, kx, ky, kz):
    nx, ny, nz = len(x), len(y), len(z)
    tx, ty, tz = np.zeros(nx + kx), np.zeros(ny + ky), np.zeros(nz + kz)
    coef = np.zeros(nx * ny * nz)
    rc = lib().oc_bspline_fit3(np.ascontiguousarray(x, dtype=np.float64), nx,
                               np.ascontiguousarray(y, dtype=np.float64), ny,
                               np.ascontiguousarray(z, dtype=np.float64), nz, kx, ky, kz, _f(f),
                               tx, ty, tz, coef)
    return rc, tx, ty, tz, coef.reshape((nx, ny, nz), order="F")


def bspline_eval3(p, tx, ty, tz, coef, kx, ky, kz):
    nx, ny, nz = coef.shape
    c = _f(coef)
    fl = C.c_int(0)
    v = lib().oc_bspline_eval3(float(p[0]), float(p[1]), float(p[2]), tx, ty, tz, nx, ny, nz,
                               kx, ky, kz, c, C.byref(fl))
    return v, fl.value


def eig_sym3_max(a):
    return lib().oc_eig_sym3_max(np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel()))


def cauchy_green(g):
    c = np.zeros(9)
    lib().oc_cauchy_green(np.ascontiguousarray(np.asarray(g, dtype=np.float64).ravel()), c)
    return c.reshape(3, 3)
