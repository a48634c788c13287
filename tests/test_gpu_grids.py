"""GPU parity for diagnostic grids other than the flow grid, through the C
ABI against the FP64 oracle:

* resolution != 0 (new_sgrid_from_sgrid, sgrid_m.f90:445-761): refined and
  coarsened grids on the periodic ABC box (where refined nodes past the last
  flow node track to the periodic ghost node, which a backward set uses until
  set_lp_bc clamps it) and on the walled double gyre;
* AUX_GRID (create_aux_grid, lcs_m.f90:227-435): six shifted particle sets
  per diagnostic and the FTLE from their differences (lcs_m.f90:95-126).

Tolerances as in test_gpu_parity: FP64 flowmaps 1e-10 relative, FTLE 1e-9
absolute.  The grid coordinates must be identical (checked in run_both), and
the integration counters count every set's particles."""
import numpy as np
import pytest

import flows as F
from scenarios import run_both
from test_gpu_parity import _check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dp():
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("dp")
    yield lib
    lib.finalize()


@pytest.mark.parametrize("res", [1, 2, -1])
@pytest.mark.parametrize("case,n", [("abc", 12), ("gyre", (33, 17))])
def test_resolution_linear_fp64(dp, res, case, n):
    stats = []
    recs, hs = run_both(dp, case, n=n, integrator=F.RK2, interpolator=F.LINEAR, T=0.6, h=0.2, dt=0.05,
                        nsteps=16, resolution=res, stats=stats)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4
    (gf, gb), (of, ob) = stats[0]
    assert (gf, gb) == (of, ob)


@pytest.mark.parametrize("res", [1, 2])
@pytest.mark.parametrize("case,n,integ", [("abc", 10, F.RK4), ("gyre_slip", (24, 12), F.RK3)])
def test_resolution_cubic_fp64(dp, res, case, n, integ):
    recs, hs = run_both(dp, case, n=n, integrator=integ, interpolator=F.CUBIC, T=0.6, h=0.2, dt=0.05,
                        nsteps=16, resolution=res)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4


@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC, F.TSE_LIMIT])
@pytest.mark.parametrize("case,n", [("abc", 10), ("gyre", (32, 16))])
def test_aux_grid_fp64(dp, interp, case, n):
    stats = []
    recs, hs = run_both(dp, case, n=n, integrator=F.RK2, interpolator=interp, T=0.6, h=0.2, dt=0.05,
                        nsteps=16, aux=True, stats=stats)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4
    (gf, gb), (of, ob) = stats[0]
    assert (gf, gb) == (of, ob)


def test_aux_grid_on_refined_grid_fp64(dp):
    recs, hs = run_both(dp, "abc", n=8, integrator=F.RK2, interpolator=F.LINEAR, T=0.6, h=0.2, dt=0.05,
                        nsteps=14, resolution=1, aux=True)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4


def test_aux_grid_ftle_of_linear_strain(dp):
    """u = -(a x, b y) on a walled square: every AUX set's composed
    displacement is exactly linear, so the AUX FTLE equals the closed form
    log(max|R^N - 1|)/T (R the RK2 amplification factor of the sub-cycle)
    inside two node layers (nearer a corner, a particle off its tracked flow
    node is projected along that node's corner-leaning wall normal)."""
    a, b, T, h, dt = 0.8, 0.3, 0.25, 0.125, 1.0 / 64
    X, Y, Z = F.cell_grid((32, 32, 1), (-1, -1, 0), (1, 1, 1))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    fl[0], fl[-1], fl[:, 0], fl[:, -1] = F.LCS_WALL, F.LCS_WALL, F.LCS_WALL, F.LCS_WALL
    dp.finalize()
    dp.init(X, Y, Z, fl)
    dp.set_option("AUX_GRID", 1)
    dp.set_option("INTEGRATOR", F.RK2)
    dp.set_option("INTERPOLATOR", F.LINEAR)
    hd = dp.diagnostic_init(F.FTLE_FWD, 0, T, h, "lin")
    for step in range(17):
        dp.update(-a * X, -b * Y, 0 * X, step * dt)
        assert dp.error == 0
    assert dp.noutput(hd) == 2
    nsub = int(dp.stats()["nsub"])
    amp = lambda z: 1 + z + z * z / 2
    cx = amp(-a * dt / nsub) ** (16 * nsub) - 1
    cy = amp(-b * dt / nsub) ** (16 * nsub) - 1
    ref = np.log(max(abs(cx), abs(cy))) / T
    f = dp.ftle(hd)
    assert np.abs(f[2:-2, 2:-2] - ref).max() < 1e-10
    dp.finalize()


@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC])
def test_refined_aux_fp32_against_fp64_oracle(interp):
    """The single-precision library on a refined grid with AUX_GRID (grid
    coordinates and shifts computed in REAL(4) as the reference's
    single-precision build does), against the FP64 oracle."""
    import libcfd2lcs_amd as L
    sp = L.Cfd2lcs("sp")
    recs, hs = run_both(sp, "abc", n=10, integrator=F.RK4, interpolator=interp, T=0.4, h=0.2, dt=0.05,
                        nsteps=10, resolution=1, aux=True)
    assert _check(recs, hs, 2e-5, 1e-3) >= 2
    sp.finalize()
