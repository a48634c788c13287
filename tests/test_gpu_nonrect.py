"""GPU parity on non-rectilinear grids, through the C ABI against the FP64
oracle: every interpolation is TSE, or TSE_LIMIT for any other choice, with
the face least-squares gradient (lp_tracking_m.f90:319-328,
gradient_m.f90:114-295); tracking, particle BCs and the 2-D projection use
the full node coordinates; composition guesses a relocated particle's node
from the coordinate extrema (comms_m.f90:1625-1688); the FTLE gradient is
least-squares (lcs_m.f90:131-133).

Cases: the walled double gyre with bent node lines (2-D) and the triply
periodic ABC box with a warp of the box's period (3-D).  Tolerances as in
test_gpu_parity (FP64 flowmaps 1e-10 relative, FTLE 1e-9 absolute)."""
import numpy as np
import pytest

import flows as F
from scenarios import make_case, run_both
from test_gpu_parity import _check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dp():
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("dp")
    yield lib
    lib.finalize()


@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC, F.TSE])
@pytest.mark.parametrize("integ", [F.RK2, F.RK4])
@pytest.mark.parametrize("case,n", [("gyre_warp", (48, 24)), ("abc_warp", 12)])
def test_non_rectilinear_fp64(dp, case, n, integ, interp):
    stats = []
    recs, hs = run_both(dp, case, n=n, integrator=integ, interpolator=interp, T=0.6, h=0.2, dt=0.05,
                        nsteps=16, stats=stats)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4
    (gf, gb), (of, ob) = stats[0]
    assert (gf, gb) == (of, ob)


def test_non_rectilinear_fp32_against_fp64_oracle():
    import libcfd2lcs_amd as L
    sp = L.Cfd2lcs("sp")
    recs, hs = run_both(sp, "abc_warp", n=16, integrator=F.RK4, interpolator=F.CUBIC, T=0.4, h=0.2, dt=0.05,
                        nsteps=10)
    assert _check(recs, hs, 2e-5, 1e-3) >= 2
    sp.finalize()


def test_non_rectilinear_refusals(dp, capfd):
    """What this build does not restate on a non-rectilinear grid fails with
    a message and the sticky error flag: the trilinear grid refinement of
    resolution != 0 (sgrid_m.f90:666-683), AUX_GRID, and the velocity
    invariants' least-squares gradient (invariants_m.f90:38-41)."""
    X, Y, Z, fl, vel = make_case("gyre_warp", (16, 8))
    for setup in ("resolution", "aux"):
        dp.finalize()
        dp.init(X, Y, Z, fl)
        if setup == "aux":
            dp.set_option("AUX_GRID", 1)
        dp.diagnostic_init(F.FTLE_FWD, 1 if setup == "resolution" else 0, 0.4, 0.2, "x")
        assert dp.error == 1
        assert "non-rectilinear" in capfd.readouterr().out
    dp.finalize()
