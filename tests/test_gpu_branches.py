"""GPU parity for the boundary flags, interpolators and options that the
smooth periodic cases do not reach: LCS_INFLOW / LCS_OUTFLOW faces,
LCS_WALL / LCS_SLIP faces and corners, interior LCS_MASK blocks,
NEAREST_NBR, QUADRATIC, TSE_LIMIT, INCOMPRESSIBLE and the first update.
The HIP path (FP64 library, through the C ABI) is compared with the FP64
oracle on the same inputs per update and per output at 1e-10 relative, and
on the constructed cases also with the closed forms that
tests/test_oracle_branches.py pins the oracle to."""
import numpy as np
import pytest

import flows as F
from scenarios import rel_err, run_both
from test_gpu_parity import _check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dp():
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("dp")
    yield lib
    lib.finalize()


@pytest.mark.parametrize("integ,interp", [
    (F.RK2, F.LINEAR), (F.RK4, F.CUBIC), (F.RK3, F.QUADRATIC), (F.RK2, F.NEAREST_NBR),
    (F.RK4, F.TSE_LIMIT), (F.EULER, F.TSE), (F.TRAPEZOIDAL, F.LINEAR)])
def test_channel_inflow_outflow_walls_mask_2d(dp, integ, interp):
    """Double gyre plus a mean flow through a channel: INFLOW at i = 1,
    OUTFLOW at i = n, WALL on the j faces and an interior MASK block.
    Forward particles stick on the outflow face and are pushed out of the
    block; the backward flowmap is zeroed on IB MASK / INFLOW / OUTFLOW."""
    recs, hs = run_both(dp, "channel", n=(64, 32), integrator=integ, interpolator=interp, T=1.0, h=0.25,
                        dt=0.05, nsteps=30)
    assert _check(recs, hs, 1e-10, 1e-9) >= 8


@pytest.mark.parametrize("integ,interp", [(F.RK4, F.LINEAR), (F.RK4, F.CUBIC), (F.RK2, F.QUADRATIC),
                                          (F.RK3, F.NEAREST_NBR), (F.RK2, F.TSE_LIMIT)])
def test_channel_inflow_outflow_slip_mask_3d(dp, integ, interp):
    recs, hs = run_both(dp, "channel3d", n=(24, 16, 12), integrator=integ, interpolator=interp, T=0.6, h=0.2,
                        dt=0.05, nsteps=16)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4


def _box(dp, case, integ, interp, diags):
    recs, hs = run_both(dp, case, integrator=integ, interpolator=interp, T=10.0, h=10.0, dt=1.0, nsteps=11,
                        diags=diags, cfl=0.25)
    assert _check(recs, hs, 1e-10, 1e-9) == 1
    out = recs[-1]["out"][hs[0]]
    return out[0]  # the GPU's composed flowmap


@pytest.mark.parametrize("integ,interp", [(F.EULER, F.LINEAR), (F.RK2, F.LINEAR), (F.RK4, F.NEAREST_NBR)])
def test_wall_box_closed_form(dp, integ, interp):
    """Uniform flow into WALL node lines: forward particles end at
    min(x0 + cT, x_n); with NEAREST_NBR the backward flowmap is 0 on the
    upstream wall and -cT elsewhere."""
    fm = _box(dp, "wall_box", integ, interp, (F.FTLE_FWD,))
    x0 = np.arange(24) + 0.5
    want = np.minimum(x0 + 5.0, 23.5) - x0
    assert np.abs(fm[0] - want[:, None, None]).max() < 1e-12
    if interp == F.NEAREST_NBR:
        fb = _box(dp, "wall_box", integ, interp, (F.FTLE_BKWD,))
        assert np.abs(fb[0][0]).max() == 0 and np.abs(fb[0][1:] + 5.0).max() < 1e-12


@pytest.mark.parametrize("integ,interp", [(F.EULER, F.LINEAR), (F.RK2, F.LINEAR), (F.RK4, F.NEAREST_NBR)])
def test_inflow_outflow_box_closed_form(dp, integ, interp):
    """Particles stick on INFLOW / OUTFLOW faces (dx = 0) and freeze at
    x = 23.25 once the hysteresis hands them to the outflow node."""
    fm = _box(dp, "io_box", integ, interp, (F.FTLE_FWD,))
    i = np.arange(1, 25)
    want = np.where(i >= 19, 23.75 - i, 5.0)
    want[0] = want[-1] = 0.0
    assert np.abs(fm[0] - want[:, None, None]).max() < 1e-12
    fb = _box(dp, "io_box", integ, interp, (F.FTLE_BKWD,))
    assert np.abs(fb[0][0]).max() == 0 and np.abs(fb[0][-1]).max() == 0


@pytest.mark.parametrize("interp", [F.LINEAR, F.NEAREST_NBR, F.CUBIC, F.TSE_LIMIT])
def test_mask_box(dp, interp):
    recs, hs = run_both(dp, "mask_box", integrator=F.RK4, interpolator=interp, T=4.0, h=4.0, dt=1.0, nsteps=5)
    assert _check(recs, hs, 1e-10, 1e-9) == 2
    X, Y, Z, fl = F.uniform_grid(12)
    fl[3:8, 3:8, 3:8] = F.LCS_MASK
    m = fl == F.LCS_MASK
    gf = recs[-1]["out"][hs[0]][0]
    gb = recs[-1]["out"][hs[1]][0]
    for d in range(3):
        assert np.abs(gb[d][m]).max() == 0
        if interp in (F.LINEAR, F.NEAREST_NBR):
            assert np.abs(gf[d][4:7, 4:7, 4:7]).max() == 0


@pytest.mark.parametrize("interp", [F.QUADRATIC, F.CUBIC])
def test_quadratic_shear_closed_form(dp, interp):
    """u = (y^2, b, 0): the splines reproduce the quadratic velocity and the
    quadratic-in-y backward flowmap, RK4 integrates y(t)^2 exactly."""
    b, T = 0.05, 0.2
    recs, hs = run_both(dp, "quad_shear", integrator=F.RK4, interpolator=interp, T=T, h=T, dt=0.05, nsteps=5)
    assert _check(recs, hs, 1e-10, 1e-9) == 2
    X, Y, Z = F.cell_grid((8, 64, 1), (0, 0, 0), (1, 1, 1))
    y = Y[0, :, 0]
    rows = slice(18, 46)
    wf = ((y + b * T) ** 3 - y ** 3) / (3 * b)
    wb = -(y ** 3 - (y - b * T) ** 3) / (3 * b)
    for hd, want, sy in ((hs[0], wf, b * T), (hs[1], wb, -b * T)):
        g = recs[-1]["out"][hd][0]
        assert np.abs(g[0][:, rows, 0] - want[None, rows]).max() < 1e-11
        assert np.abs(g[1][:, rows, 0] - sy).max() < 1e-12


@pytest.mark.parametrize("integ", [F.RK2, F.RK4])
def test_nearest_abc(dp, integ):
    """NEAREST_NBR on the periodic ABC box: forward particles change node
    along their paths, so the velocities they read depend on every tracking
    decision; the backward flowmap is -(t - t0) u(x) exactly (steady flow)."""
    recs, hs = run_both(dp, "abc", n=12, integrator=integ, interpolator=F.NEAREST_NBR, T=0.6, h=0.2, dt=0.05,
                        nsteps=16)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4
    X, Y, Z, fl = F.abc_grid(12)
    u = F.abc_velocity(X, Y, Z)
    g = recs[-1]["fm"][hs[1]][0]  # 3 steps after the output at t = 0.6
    for d in range(3):
        assert np.abs(g[d] + 0.15 * u[d]).max() < 1e-13


@pytest.mark.parametrize("case,n", [("abc", 12), ("gyre_slip", (48, 24))])
def test_quadratic_spline(dp, case, n):
    recs, hs = run_both(dp, case, n=n, integrator=F.RK4, interpolator=F.QUADRATIC, T=0.6, h=0.2, dt=0.05,
                        nsteps=16)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4


@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC])
def test_incompressible(dp, interp):
    """INCOMPRESSIBLE = max(ftle, 0): the channel's backward FTLE has
    negative values that the option clamps."""
    recs, hs = run_both(dp, "channel", n=(48, 24), integrator=F.RK2, interpolator=interp, T=0.5, h=0.25,
                        dt=0.05, nsteps=16, options=(("INCOMPRESSIBLE", 1),))
    assert _check(recs, hs, 1e-10, 1e-9) >= 4
    for r in recs:
        for hd, (gc, oc, gf, of) in r["out"].items():
            assert np.nanmin(gf) >= 0.0


def test_first_update_moves_nothing(dp):
    X, Y, Z, fl = F.double_gyre_grid(32, 16)
    dp.finalize()
    dp.init(X, Y, Z, fl)
    hf = dp.diagnostic_init(F.FTLE_FWD, 0, 1.0, 0.5, "f")
    hb = dp.diagnostic_init(F.FTLE_BKWD, 0, 1.0, 0.5, "b")
    dp.set_option("INTEGRATOR", F.RK4)
    dp.update(*F.double_gyre_velocity(X, Y, 0.3), 0.3)
    assert dp.error == 0
    st = dp.stats()
    assert st["integrations_fwd"] == 0 and st["integrations_bkwd"] == 0
    for h in (hf, hb):
        assert all(np.all(a == 0) for a in dp.flowmap(h))


def test_periodic_z_with_slip_y_quirk(dp):
    """z periodic with y walled: the ghost z coordinates follow the
    reference's periodic_shift(i,k,k,3) (comms_m.f90:750), so the z-face
    ghosts keep the opposite interior coordinate unshifted while the y-z
    edges keep their mirror values.  The oracle runs such a grid (its LINEAR
    brackets then straddle a non-monotone line; a spline fit fails); this
    build keeps one coordinate line per axis and refuses the grid with a
    message and the error flag instead of computing on it."""
    import scenarios as S
    from oracle_api import Oracle
    X, Y, Z, fl, vel = S.make_case("abc_slip_y")
    o = Oracle(X, Y, Z, fl)
    o.diagnostic_init(F.FTLE_BKWD, 0, 0.6, 0.2, "b")
    o.update(*vel(0.0), 0.0)
    o.update(*vel(0.05), 0.05)
    assert o.error == 0
    o.set_option("INTERPOLATOR", F.CUBIC)
    o.update(*vel(0.1), 0.1)
    assert o.error != 0
    o.close()
    dp.finalize()
    dp.init(X, Y, Z, fl)
    assert dp.error != 0
    dp.finalize()
    assert dp.error == 0
