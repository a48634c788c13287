"""GPU parity: the HIP path through the C ABI against the FP64 oracle on the
same inputs, per update (h-interval flowmap) and per output (composed
T-flowmap and FTLE).  FP64 tolerance 1e-10 relative (flowmaps); FTLE compared
where it is finite and above log(tiny)/T."""
import numpy as np
import pytest

import flows as F
from scenarios import rel_err, run_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dp():
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("dp")
    yield lib
    lib.finalize()


def _check(recs, handles, tol_fm, tol_ftle):
    nout = 0
    for r in recs:
        for hd, (g, o) in r["fm"].items():
            for c in range(3):
                e = rel_err(g[c], o[c])
                assert e <= tol_fm, f"step {r['step']} handle {hd} comp {c}: flowmap rel err {e:.3e}"
        for hd, (gc, oc, gf, of) in r["out"].items():
            nout += 1
            for c in range(3):
                e = rel_err(gc[c], oc[c])
                assert e <= tol_fm, f"step {r['step']} handle {hd}: composed rel err {e:.3e}"
            m = np.isfinite(of) & (of > -50)
            assert np.array_equal(np.isfinite(gf), np.isfinite(of))
            d = np.abs(gf[m] - of[m]).max() if m.any() else 0.0
            assert d <= tol_ftle, f"step {r['step']} handle {hd}: ftle abs err {d:.3e}"
    return nout


def test_abc_rk2_linear_fp64(dp):
    recs, hs = run_both(dp, "abc", n=16, integrator=F.RK2, interpolator=F.LINEAR, T=1.0, h=0.25,
                        dt=0.05, nsteps=30)
    assert _check(recs, hs, 1e-10, 1e-9) >= 8


def test_gyre_walls_rk2_linear_fp64(dp):
    recs, hs = run_both(dp, "gyre", n=(64, 32), integrator=F.RK2, interpolator=F.LINEAR, T=1.0, h=0.25,
                        dt=0.05, nsteps=30)
    assert _check(recs, hs, 1e-10, 1e-9) >= 8


@pytest.mark.parametrize("integ", [F.EULER, F.TRAPEZOIDAL, F.RK3, F.RK4])
def test_abc_integrators_linear_fp64(dp, integ):
    recs, hs = run_both(dp, "abc", n=12, integrator=integ, interpolator=F.LINEAR, T=0.6, h=0.2,
                        dt=0.05, nsteps=16)
    _check(recs, hs, 1e-10, 1e-9)


def test_abc_rk4_cubic_fp64(dp):
    recs, hs = run_both(dp, "abc", n=12, integrator=F.RK4, interpolator=F.CUBIC, T=0.6, h=0.2,
                        dt=0.05, nsteps=16)
    _check(recs, hs, 1e-10, 1e-9)


def test_gyre_slip_rk3_cubic_fp64(dp):
    recs, hs = run_both(dp, "gyre_slip", n=(32, 16), integrator=F.RK3, interpolator=F.CUBIC, T=0.6, h=0.2,
                        dt=0.05, nsteps=16)
    _check(recs, hs, 1e-10, 1e-9)


def test_abc_rk4_cubic_fp64_multibrick(dp):
    """36^3 spans several 8x4x4 / 16x4x4 bricks, including partial ones, so
    the tiled semi-Lagrangian kernel's halo and fallback paths are covered."""
    recs, hs = run_both(dp, "abc", n=36, integrator=F.RK4, interpolator=F.CUBIC, T=0.2, h=0.1,
                        dt=0.05, nsteps=5, every_step=2)
    assert _check(recs, hs, 1e-10, 1e-9) >= 2


@pytest.fixture(scope="module")
def sp():
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("sp")
    yield lib
    lib.finalize()


@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC])
def test_abc_fp32_against_fp64_oracle(sp, interp):
    """FP32 library (lcsdata_t = float, FP32 compute) against the FP64
    oracle: flowmaps within 2e-5 relative, FTLE within 1e-3 absolute."""
    recs, hs = run_both(sp, "abc", n=20, integrator=F.RK4, interpolator=interp, T=0.4, h=0.2,
                        dt=0.05, nsteps=10, every_step=3)
    assert _check(recs, hs, 2e-5, 1e-3) >= 2
