"""Pins the torch tangent-linear reference (tests/tlm_reference.py) to the FP64
oracle: its forward FTLE equals the oracle's backward FTLE with h = T on a
periodic 2-D grid (LINEAR, RK2), and its tangent agrees with central
differences of itself, so that it can stand as the FP64 tangent-linear model
the GPU adjoint is checked against."""
import numpy as np
import pytest

import flows as F
from oracle_api import Oracle

torch = pytest.importorskip("torch")
import tlm_reference as TR  # noqa: E402


def periodic_case(n0=24, n1=16, S=5, dt=0.08):
    L = 2 * np.pi
    X, Y, Z = F.cell_grid((n0, n1, 1), (0, 0, 0), (L, L, 1.0))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    times = [dt * s for s in range(S)]
    us, vs = [], []
    for t in times:
        m = 1.0 + 0.3 * np.sin(2.0 * t)
        us.append(np.sin(X) * np.cos(Y) * m + 0.1 * np.cos(Y + t))
        vs.append(-np.cos(X) * np.sin(Y) * m)
    return X, Y, Z, fl, times, us, vs


def test_torch_forward_matches_oracle():
    X, Y, Z, fl, times, us, vs = periodic_case()
    T = times[-1] - times[0]
    o = Oracle(X, Y, Z, fl)
    hd = o.diagnostic_init(F.FTLE_BKWD, 0, T, T, "b")
    o.set_option("INTEGRATOR", F.RK2)
    o.set_option("INTERPOLATOR", F.LINEAR)
    for u, v, t in zip(us, vs, times):
        o.update(u, v, np.zeros_like(u), t)
    assert o.error == 0 and o.noutput(hd) == 1
    f_or = o.ftle(hd)[:, :, 0].T  # (n1, n0)
    o.close()
    ut = torch.tensor(np.stack([a[:, :, 0].T for a in us]), dtype=torch.float64)
    vt = torch.tensor(np.stack([a[:, :, 0].T for a in vs]), dtype=torch.float64)
    x = torch.tensor(X[:, 0, 0], dtype=torch.float64)
    y = torch.tensor(Y[0, :, 0], dtype=torch.float64)
    f_t, _ = TR.bkwd_ftle_2d(ut, vt, times, x, y)
    assert np.abs(f_t.numpy() - f_or).max() <= 1e-11


def test_tangent_matches_central_differences():
    X, Y, Z, fl, times, us, vs = periodic_case()
    ut = torch.tensor(np.stack([a[:, :, 0].T for a in us]), dtype=torch.float64)
    vt = torch.tensor(np.stack([a[:, :, 0].T for a in vs]), dtype=torch.float64)
    x = torch.tensor(X[:, 0, 0], dtype=torch.float64)
    y = torch.tensor(Y[0, :, 0], dtype=torch.float64)
    _, subs = TR.bkwd_ftle_2d(ut, vt, times, x, y)
    g = torch.Generator().manual_seed(1)
    w = torch.randn(ut.shape[1:], generator=g, dtype=torch.float64)
    du = torch.randn(ut.shape, generator=g, dtype=torch.float64)
    dv = torch.randn(vt.shape, generator=g, dtype=torch.float64)

    def L(a, b):
        return (w * TR.bkwd_ftle_2d(a, b, times, x, y, nsub=subs)[0]).sum()

    _, jv = torch.func.jvp(L, (ut, vt), (du, dv))
    h = 1e-6
    fd = (L(ut + h * du, vt + h * dv) - L(ut - h * du, vt - h * dv)) / (2 * h)
    assert abs(float(jv) - float(fd)) <= 1e-5 * abs(float(fd))
