"""Analytic grids and velocity fields for the configurations the reference's
example drivers use (examples/abc_flow_c.c:164-228, examples/double_gyre_c.c:164-234)
and for the closed-form known-answer tests.  Arrays are (ni, nj, nk), Fortran
order when flattened (i fastest), as the C API expects."""
import numpy as np

LCS_INTERNAL, LCS_INFLOW, LCS_OUTFLOW, LCS_WALL, LCS_SLIP, LCS_2D, LCS_MASK = range(7)
FTLE_FWD, FTLE_BKWD, LP_TRACER = 0, 1, 2
EULER, TRAPEZOIDAL, RK2, RK3, RK4 = 0, 1, 2, 3, 4
NEAREST_NBR, LINEAR, QUADRATIC, CUBIC, TSE, TSE_LIMIT = 0, 1, 2, 3, 5, 6


def cell_grid(n, lo, hi):
    """Cell-centred uniform grid: x_i = lo + (i + 1/2) (hi - lo) / n."""
    axes = []
    for d in range(3):
        h = (hi[d] - lo[d]) / max(n[d], 1)
        axes.append(lo[d] + 0.5 * h + np.arange(n[d]) * h)
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    return X, Y, Z


def abc_grid(n):
    L = 2.0 * np.pi
    X, Y, Z = cell_grid((n, n, n) if np.isscalar(n) else n, (0, 0, 0), (L, L, L))
    flag = np.full(X.shape, LCS_INTERNAL, dtype=np.int32)
    return X, Y, Z, flag


def abc_velocity(X, Y, Z, A=np.sqrt(3.0), B=np.sqrt(2.0), Cc=1.0):
    u = A * np.sin(Z) + Cc * np.cos(Y)
    v = B * np.sin(X) + A * np.cos(Z)
    w = Cc * np.sin(Y) + B * np.cos(X)
    return u, v, w


def double_gyre_grid(ni, nj, walls=True):
    """[0,2]x[0,1] cell-centred, one plane in z; LCS_WALL on the outermost
    i and j node lines."""
    X, Y, Z = cell_grid((ni, nj, 1), (0.0, 0.0, 0.0), (2.0, 1.0, 1.0))
    flag = np.full(X.shape, LCS_INTERNAL, dtype=np.int32)
    if walls:
        flag[0, :, :] = LCS_WALL
        flag[-1, :, :] = LCS_WALL
        flag[:, 0, :] = LCS_WALL
        flag[:, -1, :] = LCS_WALL
    return X, Y, Z, flag


def double_gyre_velocity(X, Y, t, A=0.1, eps=0.1, omega=2 * np.pi / 10):
    a = eps * np.sin(omega * t)
    b = 1.0 - 2.0 * eps * np.sin(omega * t)
    f = a * X ** 2 + b * X
    dfdx = 2.0 * a * X + b
    u = -np.pi * A * np.sin(np.pi * f) * np.cos(np.pi * Y)
    v = np.pi * A * np.cos(np.pi * f) * np.sin(np.pi * Y) * dfdx
    w = np.zeros_like(X)
    return u, v, w
