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


@pytest.mark.parametrize("interp", [F.TSE, F.TSE_LIMIT])
@pytest.mark.parametrize("case,n", [("abc", 12), ("gyre", (48, 24))])
def test_tse_fp64(dp, interp, case, n):
    """Taylor-series interpolation (lp_tracking_m.f90:526-654) on the periodic
    ABC box and on the walled double gyre."""
    recs, hs = run_both(dp, case, n=n, integrator=F.RK2, interpolator=interp, T=0.6, h=0.2, dt=0.05,
                        nsteps=16)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4


def test_abc_rk4_cubic_fp64_multibrick(dp):
    """36^3: more than one workgroup per grid row and a partial last block."""
    recs, hs = run_both(dp, "abc", n=36, integrator=F.RK4, interpolator=F.CUBIC, T=0.2, h=0.1,
                        dt=0.05, nsteps=5, every_step=2)
    assert _check(recs, hs, 1e-10, 1e-9) >= 2


@pytest.fixture(scope="module")
def sp():
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("sp")
    yield lib
    lib.finalize()


@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC, F.TSE_LIMIT])
def test_abc_fp32_against_fp64_oracle(sp, interp):
    """FP32 library (lcsdata_t = float, FP32 compute) against the FP64
    oracle: flowmaps within 2e-5 relative, FTLE within 1e-3 absolute."""
    recs, hs = run_both(sp, "abc", n=20, integrator=F.RK4, interpolator=interp, T=0.4, h=0.2,
                        dt=0.05, nsteps=10, every_step=3)
    assert _check(recs, hs, 2e-5, 1e-3) >= 2


def _sl_run(lib, tiled, integ, interp, shape, cfl, walls, nsteps=8, opts=None):
    import libcfd2lcs_amd as L
    X, Y, Z = F.cell_grid(shape, (0, 0, 0), (2 * np.pi, 2 * np.pi, 2 * np.pi))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    if walls:
        fl[0], fl[-1] = F.LCS_WALL, F.LCS_WALL
        fl[:, :, 0], fl[:, :, -1] = F.LCS_SLIP, F.LCS_SLIP
    lib.finalize()
    lib.init(X, Y, Z, fl)
    # storage options take effect at a reset: before the diagnostic
    for k, v in (opts or {}).items():
        lib.set_option(k, v)
    hd = lib.diagnostic_init(L.FTLE_BKWD, 0, 0.3, 0.15, "b")
    lib.set_option("INTEGRATOR", integ)
    lib.set_option("INTERPOLATOR", interp)
    lib.set_option("SL_TILED", L.LCS_TRUE if tiled else L.LCS_FALSE)
    lib.set_param("CFL", cfl)
    u, v, w = F.abc_velocity(X, Y, Z)
    fms, outs = [], []
    for s in range(nsteps):
        lib.update(u, v, w, 0.05 * s)
        assert lib.error == 0
        fms.append(lib.flowmap(hd))
        if lib.noutput(hd) > len(outs):
            outs.append((lib.composed(hd), lib.ftle(hd)))
    return fms, outs


@pytest.mark.parametrize("integ,interp,shape,cfl,walls", [
    (F.RK4, F.CUBIC, (20, 18, 13), 0.5, False),     # partial bricks in every direction
    (F.RK3, F.CUBIC, (24, 16, 12), 0.5, True),      # walls and slip faces
    (F.RK2, F.QUADRATIC, (17, 19, 9), 0.5, False),
    (F.RK4, F.CUBIC, (16, 16, 16), 2.0, False),     # displacements beyond the tile halo
])
def test_sl_lds_tiles_are_bit_exact(sp, integ, interp, shape, cfl, walls):
    """The LDS-tiled semi-Lagrangian step (FP32, 3-D, B-spline) gives the
    untiled gather kernel's results bit for bit: every flowmap and FTLE."""
    a_fm, a_out = _sl_run(sp, True, integ, interp, shape, cfl, walls)
    b_fm, b_out = _sl_run(sp, False, integ, interp, shape, cfl, walls)
    assert len(a_out) == len(b_out) >= 2
    for x, y in zip(a_fm, b_fm):
        for c in range(3):
            assert np.array_equal(x[c], y[c])
    for (ca, fa), (cb, fb) in zip(a_out, b_out):
        for c in range(3):
            assert np.array_equal(ca[c], cb[c])
        assert np.array_equal(fa, fb, equal_nan=True)


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
@pytest.mark.parametrize("integ,shape,cfl,walls", [
    (F.RK4, (40, 18, 13), 0.5, False),     # partial bricks in every direction
    (F.RK3, (32, 16, 12), 0.5, True),      # walls and slip faces
    (F.RK4, (16, 16, 16), 2.0, False),     # displacements beyond the tile halo
])
def test_sl_lds_tiles_are_bit_exact_fp64_compute(dp, integ, shape, cfl, walls, storage):
    """Under FP64 compute, with FP64 fields or with FP32 velocity and flowmap
    storage, the LDS-tiled semi-Lagrangian step gives the gather kernel's
    results bit for bit."""
    import libcfd2lcs_amd as L
    opts = {"VELOCITY_STORAGE": L.LCS_FP32, "FLOWMAP_STORAGE": L.LCS_FP32} if storage == "fp32" else {}
    a_fm, a_out = _sl_run(dp, True, integ, F.CUBIC, shape, cfl, walls, opts=opts)
    b_fm, b_out = _sl_run(dp, False, integ, F.CUBIC, shape, cfl, walls, opts=opts)
    assert len(a_out) == len(b_out) >= 2
    for x, y in zip(a_fm, b_fm):
        for c in range(3):
            assert np.array_equal(x[c], y[c])
    for (ca, fa), (cb, fb) in zip(a_out, b_out):
        for c in range(3):
            assert np.array_equal(ca[c], cb[c])
        assert np.array_equal(fa, fb, equal_nan=True)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_sl_lds_tiles_are_bit_exact_at_256(sp, dp, prec):
    """At the headline size (ABC 256^3, RK4 + cubic, about 17 sub-cycles per
    update) the LDS-brick step and the gather kernel give the same flowmap
    bit for bit after every update, in FP32 and in FP64: a size-independent
    check of the brick kernel where the oracle would take too long."""
    lib = sp if prec == "fp32" else dp
    a_fm, _ = _sl_run(lib, True, F.RK4, F.CUBIC, (256, 256, 256), 0.5, False, nsteps=3)
    b_fm, _ = _sl_run(lib, False, F.RK4, F.CUBIC, (256, 256, 256), 0.5, False, nsteps=3)
    assert len(a_fm) == len(b_fm) == 3
    moved = 0.0
    for x, y in zip(a_fm, b_fm):
        for c in range(3):
            assert np.array_equal(x[c], y[c])
            moved = max(moved, float(np.abs(x[c]).max()))
    assert moved > 0.01  # the flowmap did move


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_uniform_flow_closed_form_at_256(sp, dp, prec):
    """Periodic box at the headline size (256^3), uniform velocity U, RK4 +
    cubic B-spline, both directions: every forward particle and every
    backward departure point moves by exactly U t, periodic wraps included,
    so the composed T-flowmaps at a full ring are +U T (FWD) and -U T (BKWD)
    at every node, and the displacement gradient, hence sigma, vanishes to
    rounding (FTLE far below any resolved value)."""
    import libcfd2lcs_amd as L
    lib = sp if prec == "fp32" else dp
    # times exact in FP32 too, so that both libraries trigger the same outputs
    n, T, h, dt = 256, 0.25, 0.125, 0.0625
    X, Y, Z, fl = F.abc_grid(n)
    lib.finalize()
    lib.init(X, Y, Z, fl)
    hf = lib.diagnostic_init(L.FTLE_FWD, 0, T, h, "fwd")
    hb = lib.diagnostic_init(L.FTLE_BKWD, 0, T, h, "bkwd")
    lib.set_option("INTEGRATOR", L.RK4)
    lib.set_option("INTERPOLATOR", L.CUBIC)
    U = (0.3, -0.2, 0.1)
    vel = [np.full(X.shape, c) for c in U]
    for s in range(10):  # outputs at t = 1/8, 1/4, 3/8, 1/2; the last composes a full ring
        lib.update(*vel, dt * s)
        assert lib.error == 0
    assert lib.noutput(hf) == lib.noutput(hb) == 4
    tol = 1e-12 if prec == "fp64" else 2e-6
    for hd, sign in ((hf, 1.0), (hb, -1.0)):
        comp = lib.composed(hd)
        for c in range(3):
            err = float(np.abs(np.asarray(comp[c], dtype=np.float64) - sign * U[c] * T).max())
            assert err <= tol, f"handle {hd} component {c}: {err:.3e}"
        ftle = np.asarray(lib.ftle(hd), dtype=np.float64)
        # log(sigma)/T with sigma below 1e-9 (FP64) or 1e-4 (FP32)
        assert np.all(ftle <= np.log(1e-9 if prec == "fp64" else 1e-4) / T)


def _fwd_run(lib, sort_every, n=24, nsteps=14):
    import libcfd2lcs_amd as L
    X, Y, Z, fl = F.abc_grid(n)
    lib.finalize()
    lib.init(X, Y, Z, fl)
    hd = lib.diagnostic_init(L.FTLE_FWD, 0, 0.4, 0.2, "fwd")
    lib.set_option("INTEGRATOR", L.RK4)
    lib.set_option("INTERPOLATOR", L.CUBIC)
    lib.set_option("PARTICLE_SORT", sort_every)
    u, v, w = F.abc_velocity(X, Y, Z)
    fms, outs = [], []
    for s in range(nsteps):
        lib.update(u, v, w, 0.05 * s)
        assert lib.error == 0
        fms.append(lib.flowmap(hd))
        if lib.noutput(hd) > len(outs):
            outs.append((lib.composed(hd), lib.ftle(hd)))
    return fms, outs


@pytest.mark.parametrize("every", [1, 3])
def test_particle_sort_is_bit_exact(dp, every):
    """Re-ordering the forward particle set by tracked node changes only the
    storage order: flowmaps, composed flowmaps and FTLE are bit-identical to
    the run that keeps origin order."""
    a_fm, a_out = _fwd_run(dp, 0)
    b_fm, b_out = _fwd_run(dp, every)
    assert len(a_out) == len(b_out) >= 2
    for x, y in zip(a_fm, b_fm):
        for c in range(3):
            assert np.array_equal(x[c], y[c])
    for (xc, xf), (yc, yf) in zip(a_out, b_out):
        for c in range(3):
            assert np.array_equal(xc[c], yc[c])
        assert np.array_equal(xf, yf, equal_nan=True)


def _vs_oracle(lib, opts, interp, n=20, T=0.4, h=0.2, nsteps=10):
    """Max relative flowmap error (composed and per-step) and max absolute
    FTLE error of the library against the FP64 oracle."""
    from oracle_api import Oracle
    X, Y, Z, fl = F.abc_grid(n)
    lib.finalize()
    lib.init(X, Y, Z, fl)
    o = Oracle(X, Y, Z, fl)
    # storage options that take effect at a reset apply from the first one
    for k, v in opts.items():
        lib.set_option(k, v)
    hs = [lib.diagnostic_init(d, 0, T, h, "x") for d in (F.FTLE_FWD, F.FTLE_BKWD)]
    for d in (F.FTLE_FWD, F.FTLE_BKWD):
        o.diagnostic_init(d, 0, T, h, "x")
    for k, v in (("INTEGRATOR", F.RK4), ("INTERPOLATOR", interp)):
        lib.set_option(k, v)
        o.set_option(k, v)
    for k, v in opts.items():
        lib.set_option(k, v)
    u, v, w = F.abc_velocity(X, Y, Z)
    efm, eftle = 0.0, 0.0
    for s in range(nsteps):
        lib.update(u, v, w, 0.05 * s)
        o.update(u, v, w, 0.05 * s)
        assert lib.error == 0 and o.error == 0
        for hd in hs:
            for c in range(3):
                efm = max(efm, rel_err(lib.flowmap(hd)[c], o.flowmap(hd)[c]))
    for hd in hs:
        assert lib.noutput(hd) == o.noutput(hd) >= 2
        for c in range(3):
            efm = max(efm, rel_err(lib.composed(hd)[c], o.composed(hd)[c]))
        eftle = max(eftle, float(np.abs(lib.ftle(hd) - o.ftle(hd)).max()))
    o.close()
    return efm, eftle


@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC])
@pytest.mark.parametrize("storage,tol_fm,tol_ftle", [("FP16", 1e-3, 1e-2), ("BF16", 8e-3, 8e-2)])
def test_16bit_velocity_storage_against_fp64_oracle(sp, interp, storage, tol_fm, tol_ftle):
    """FP32 advection with the gathered stage velocities stored in FP16 or
    BF16 (unit roundoff 2^-11 and 2^-8): flowmaps and FTLE stay within a few
    storage roundoffs of the FP64 oracle, and the 16-bit run is measurably
    different from the FP32-storage run (the 16-bit path is the one that
    ran)."""
    import libcfd2lcs_amd as L
    val = L.LCS_FP16 if storage == "FP16" else L.LCS_BF16
    e16, f16 = _vs_oracle(sp, {"VELOCITY_STORAGE": val}, interp)
    e32, f32 = _vs_oracle(sp, {}, interp)
    print(f"{storage} interp {interp}: flowmap {e16:.3e} (fp32 {e32:.3e}), ftle {f16:.3e} (fp32 {f32:.3e})")
    assert e16 <= tol_fm and f16 <= tol_ftle
    assert e16 > 2 * e32


@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC])
def test_compensated_positions_against_fp64_oracle(sp, interp):
    """Two-sum position/displacement updates of the forward particles keep
    the FP32 run at least as close to the FP64 oracle as plain FP32."""
    import libcfd2lcs_amd as L
    ec, fc = _vs_oracle(sp, {"COMPENSATED": L.LCS_TRUE}, interp)
    ep, fp = _vs_oracle(sp, {}, interp)
    print(f"compensated interp {interp}: flowmap {ec:.3e} (plain {ep:.3e}), ftle {fc:.3e} (plain {fp:.3e})")
    assert ec <= 2e-5 and fc <= 1e-3
    assert ec <= 1.5 * ep


@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC])
@pytest.mark.parametrize("which", ["velocity", "flowmap", "both"])
def test_fp32_storage_under_fp64_compute_against_oracle(dp, interp, which):
    """FP64 compute with FP32 storage of the gathered stage velocities
    (VELOCITY_STORAGE = LCS_FP32) and/or of the backward flowmap and its
    coefficients (FLOWMAP_STORAGE = LCS_FP32): flowmaps and FTLE stay within
    a few FP32 roundoffs of the FP64 oracle, and differ measurably from the
    all-FP64 run (the FP32 storage path is the one that ran)."""
    import libcfd2lcs_amd as L
    opts = {}
    if which in ("velocity", "both"):
        opts["VELOCITY_STORAGE"] = L.LCS_FP32
    if which in ("flowmap", "both"):
        opts["FLOWMAP_STORAGE"] = L.LCS_FP32
    e32, f32 = _vs_oracle(dp, opts, interp)
    e64, f64 = _vs_oracle(dp, {}, interp)
    print(f"{which} interp {interp}: flowmap {e32:.3e} (fp64 {e64:.3e}), ftle {f32:.3e} (fp64 {f64:.3e})")
    assert e32 <= 2e-6 and f32 <= 2e-5
    assert e32 > 100 * e64


def test_16bit_storage_needs_fp32_compute(dp):
    import libcfd2lcs_amd as L
    X, Y, Z, fl = F.abc_grid(8)
    dp.finalize()
    dp.init(X, Y, Z, fl)
    dp.diagnostic_init(L.FTLE_BKWD, 0, 0.2, 0.1, "b")
    dp.set_option("VELOCITY_STORAGE", L.LCS_BF16)
    u, v, w = F.abc_velocity(X, Y, Z)
    dp.update(u, v, w, 0.0)
    assert dp.error != 0
    dp.finalize()
    assert dp.error == 0


def test_double_gyre_512x256_walls_rk2_linear_fp64(dp):
    """The reference's CPU configuration at its full grid size: double gyre
    512 x 256 x 1 cell-centred on [0,2] x [0,1] with walls, dt = 0.01, RK2 +
    LINEAR, FWD and BKWD, on a window shortened to T = 1 (h = 0.5) so that the
    FP64 oracle finishes in seconds."""
    recs, hs = run_both(dp, "gyre", n=(512, 256), integrator=F.RK2, interpolator=F.LINEAR, T=1.0, h=0.5,
                        dt=0.01, nsteps=102, every_step=25)
    assert _check(recs, hs, 1e-10, 1e-9) >= 4


def test_double_gyre_512x256_full_window_fwd_fp64(dp):
    """The reference's CPU configuration over its whole window: double gyre
    512 x 256 x 1, walls, dt = 0.01, RK2 + LINEAR, forward FTLE at T = 15
    with h = 1, i.e. a 15-slot ring.  1501 updates reach t = 15, whose output
    is the first to compose all 15 sub-intervals; every output on the way
    composes the slots written so far (flowmap_m.f90:253-268).  Output
    counts and integration counters agree exactly, the FTLE to 1e-9.

    The per-slot flowmaps agree to ~1e-14 (FP64 rounding: the GPU forms
    displacements from stage increments and interpolates from split
    coordinates, the oracle from positions), and composing k slots carries
    that difference through the stretching of k time units, which doubles
    it per slot on this flow (1.3e-14 with one slot, 1.1e-10 with 14).  So
    the composed maps are held to 1e-10 through 13 slots and, beyond, to at
    most 2.5x growth per slot from there."""
    stats = []
    recs, hs = run_both(dp, "gyre", n=(512, 256), integrator=F.RK2, interpolator=F.LINEAR, T=15.0, h=1.0,
                        dt=0.01, nsteps=1501, every_step=500, diags=(F.FTLE_FWD,), stats=stats)
    (gf, gb), (of_, ob) = stats[0]
    assert gf == of_ and gb == ob == 0
    assert dp.noutput(hs[0]) == 15
    outs = [r["out"][hs[0]] for r in recs if hs[0] in r["out"]]
    assert len(outs) == 15
    errs = []
    for k, (gc, oc, gftle, oftle) in enumerate(outs):
        errs.append(max(rel_err(gc[c], oc[c]) for c in range(3)))
        m = np.isfinite(oftle) & (oftle > -50)
        fe = np.abs(gftle[m] - oftle[m]).max()
        print(f"output {k + 1:2d} ({k + 1:2d} slots): composed flowmap rel {errs[-1]:.3e}, FTLE abs {fe:.3e}")
        assert fe <= 1e-9
    assert errs[0] <= 2e-14
    assert max(errs[:13]) <= 1e-10
    for k in (13, 14):
        assert errs[k] <= 2.5 * errs[k - 1]


@pytest.mark.parametrize("k", [17, -20])
@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC])
def test_fp16_storage_is_scale_invariant(sp, k, interp):
    """FP16 stage fields hold u / 2^e with 2^e set per update from the
    largest |u|.  The velocity times a = 2^k over times divided by a traces
    the same paths: with a = 2^17 (|u| ~ 3.6e5, past FP16's 65504) and
    a = 2^-20 (|u| ~ 2.6e-6, FP16 subnormal) the flowmaps are bit-identical
    to the a = 1 run and the FTLE is a times it."""
    import libcfd2lcs_amd as L
    X, Y, Z, fl = F.abc_grid(16)
    u0 = F.abc_velocity(X, Y, Z)
    res = []
    for a in (1.0, 2.0 ** k):
        sp.finalize()
        sp.init(X, Y, Z, fl)
        hs = [sp.diagnostic_init(d, 0, 0.4 / a, 0.2 / a, "x") for d in (F.FTLE_FWD, F.FTLE_BKWD)]
        for name, val in (("INTEGRATOR", F.RK4), ("INTERPOLATOR", interp), ("VELOCITY_STORAGE", L.LCS_FP16)):
            sp.set_option(name, val)
        for s in range(10):
            sp.update(*(a * c for c in u0), np.float32(0.05 * s) / np.float32(a))
            assert sp.error == 0
        res.append([(sp.composed(h), sp.ftle(h), sp.flowmap(h)) for h in hs])
        assert all(sp.noutput(h) == 2 for h in hs)
    a = 2.0 ** k
    for (c1, f1, m1), (c2, f2, m2) in zip(*res):
        for d in range(3):
            assert np.array_equal(c1[d], c2[d]) and np.array_equal(m1[d], m2[d])
        m = np.isfinite(f1) & (f1 > -50)
        assert np.allclose(f2[m] / a, f1[m], rtol=1e-6, atol=0)


@pytest.mark.parametrize("case,n,interp", [("abc", 24, F.CUBIC), ("gyre", (48, 24), F.LINEAR),
                                           ("abc_warp", 12, F.TSE_LIMIT), ("channel", (48, 24), F.QUADRATIC)])
def test_compose_reordering_is_bit_exact(dp, case, n, interp):
    """The composition one slot per launch with the particles re-ordered by
    brick between slots (COMPOSE_SORT = 1, the default; 2: every other slot)
    gives bit for bit the single-launch composition (COMPOSE_SORT = 0):
    each particle's arithmetic does not depend on the order."""
    import scenarios as S
    out = {}
    for every in (0, 1, 2):
        X, Y, Z, fl, vel = S.make_case(case, n)
        dp.finalize()
        dp.init(X, Y, Z, fl)
        hs = [dp.diagnostic_init(d, 0, 0.6, 0.1, "x") for d in (F.FTLE_FWD, F.FTLE_BKWD)]
        for k, v in (("INTEGRATOR", F.RK4), ("INTERPOLATOR", interp), ("COMPOSE_SORT", every)):
            dp.set_option(k, v)
        res = []
        for s in range(16):
            dp.update(*vel(0.05 * s), 0.05 * s)
            assert dp.error == 0
            if s in (7, 15):
                res.append([(dp.composed(h), dp.ftle(h)) for h in hs])
        out[every] = res
    for every in (1, 2):
        for a, b in zip(out[0], out[every]):
            for (ca, fa), (cb, fb) in zip(a, b):
                for d in range(3):
                    assert np.array_equal(ca[d], cb[d])
                assert np.array_equal(fa, fb, equal_nan=True)


@pytest.mark.parametrize("variant,case,n,interp,storage", [
    ("dp", "abc", 20, F.CUBIC, None), ("sp", "abc", 24, F.CUBIC, None), ("sp", "gyre_slip", (48, 24), F.QUADRATIC, None),
    ("sp", "abc", 24, F.CUBIC, "FP16"), ("sp", "channel3d", (24, 16, 12), F.CUBIC, None)])
def test_spline_table_staging_is_bit_exact(variant, case, n, interp, storage):
    """The B-spline steps read their basis tables and knots from a
    workgroup's LDS copy (SPLINE_LDS = 1, the default) or from global memory
    (0): the same values, so the flowmaps agree bit for bit."""
    import libcfd2lcs_amd as L
    import scenarios as S
    lib = L.Cfd2lcs(variant)
    out = []
    for lds in (0, 1):
        X, Y, Z, fl, vel = S.make_case(case, n)
        lib.finalize()
        lib.init(X, Y, Z, fl)
        hs = [lib.diagnostic_init(d, 0, 0.6, 0.2, "x") for d in (F.FTLE_FWD, F.FTLE_BKWD)]
        opts = [("INTEGRATOR", F.RK4), ("INTERPOLATOR", interp), ("SPLINE_LDS", lds)]
        if storage:
            opts.append(("VELOCITY_STORAGE", L.LCS_FP16))
        for k, v in opts:
            lib.set_option(k, v)
        res = []
        for s in range(10):
            lib.update(*vel(0.05 * s), 0.05 * s)
            assert lib.error == 0
            res.append([lib.flowmap(h) for h in hs])
        res.append([lib.composed(h) for h in hs])
        out.append(res)
    lib.finalize()
    for a, b in zip(*out):
        for fa, fb in zip(a, b):
            for d in range(3):
                assert np.array_equal(fa[d], fb[d])
