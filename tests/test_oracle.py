"""Pins the FP64 oracle (oracle/oracle.c) to what the mathematics fixes, since
the reference ships no tests, golden vectors or fixtures for this path:

* B-spline knots/fit/evaluation against scipy.interpolate with the same knot
  vector (dbknot rule, bspline_sub_module.f90:1737-1785) and against
  polynomial reproduction;
* the 3x3 eigen-solve and C = G^T G against numpy;
* integrators, the semi-Lagrangian step, forward particles, composition and
  FTLE against closed forms for uniform and linear-strain flows, including the
  reference's RK3 (k3 at xn + 2 k2, lp_motion_m.f90:239-247), whose
  amplification factor differs from classical RK3;
* grid set-up (ghost/fake coordinates, periodic shifts, normals, 1/delta).
"""
import numpy as np
import pytest
from scipy.interpolate import BSpline, NdBSpline, make_interp_spline

import flows as F
import oracle_api as OA
from oracle_api import Oracle


@pytest.fixture(scope="module", autouse=True)
def _built():
    OA.lib()


# ---------------------------------------------------------------- B-splines
def dbknot_rule(x, k):
    """The knot rule stated in bspline_sub_module.f90:1737-1785."""
    n = len(x)
    t = np.empty(n + k)
    t[:k] = x[0]
    t[n:] = x[-1] + 0.1 * (x[-1] - x[-2])
    if k % 2:
        for j in range(k + 1, n + 1):
            m = j + (k - 1) // 2 - k
            t[j - 1] = 0.5 * (x[m - 1] + x[m])
    else:
        for j in range(k + 1, n + 1):
            t[j - 1] = x[j + k // 2 - k - 1]
    return t


@pytest.mark.parametrize("k", [2, 3, 4])
def test_knots_match_rule(k):
    x = np.cumsum(np.random.default_rng(1).uniform(0.5, 1.5, 11))
    np.testing.assert_array_equal(OA.bspline_knots(x, k), dbknot_rule(x, k))


@pytest.mark.parametrize("k", [3, 4])
def test_fit3_matches_scipy(k):
    rng = np.random.default_rng(7)
    x = np.cumsum(rng.uniform(0.5, 1.5, 9))
    y = np.cumsum(rng.uniform(0.5, 1.5, 8))
    z = np.cumsum(rng.uniform(0.5, 1.5, 7))
    f = rng.normal(size=(9, 8, 7))
    rc, tx, ty, tz, coef = OA.bspline_fit3(x, y, z, f, k, k, k)
    assert rc == 0
    # scipy: interpolate along each axis with the same knots
    c = f.copy()
    for ax, (pts, t) in enumerate(((x, tx), (y, ty), (z, tz))):
        c = np.moveaxis(make_interp_spline(pts, c, k=k - 1, t=t, axis=ax).c, 0, ax)
    np.testing.assert_allclose(coef, c, rtol=0, atol=1e-11 * np.abs(c).max())
    nd = NdBSpline((tx, ty, tz), coef, k - 1)
    pts = np.stack([rng.uniform(x[0], x[-1], 200), rng.uniform(y[0], y[-1], 200),
                    rng.uniform(z[0], z[-1], 200)], axis=1)
    ref = nd(pts)
    got = np.array([OA.bspline_eval3(p, tx, ty, tz, coef, k, k, k)[0] for p in pts])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
    # the data are reproduced at the abscissae
    for (i, j, l) in [(0, 0, 0), (8, 7, 6), (3, 2, 5)]:
        v, fl = OA.bspline_eval3((x[i], y[j], z[l]), tx, ty, tz, coef, k, k, k)
        assert fl == 0
        assert abs(v - f[i, j, l]) < 1e-11


def test_cubic_reproduces_cubics_and_out_of_range_is_zero():
    rng = np.random.default_rng(3)
    x = np.linspace(-1.0, 1.2, 12)
    y = np.linspace(0.0, 2.0, 10)
    z = np.linspace(0.5, 1.0, 9)
    P = lambda a, b, c: 1.0 + a - 2 * a ** 3 + 0.5 * b ** 2 * c + c ** 3 - a * b * c
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    rc, tx, ty, tz, coef = OA.bspline_fit3(x, y, z, P(X, Y, Z), 4, 4, 4)
    assert rc == 0
    for _ in range(100):
        p = (rng.uniform(x[0], x[-1]), rng.uniform(y[0], y[-1]), rng.uniform(z[0], z[-1]))
        v, fl = OA.bspline_eval3(p, tx, ty, tz, coef, 4, 4, 4)
        assert fl == 0 and abs(v - P(*p)) < 1e-11
    # db3val: 0 outside [t(1), t(n+k)] (bspline_sub_module.f90:612-629)
    v, fl = OA.bspline_eval3((x[0] - 1e-3, y[3], z[3]), tx, ty, tz, coef, 4, 4, 4)
    assert v == 0.0 and fl == 1
    # ... and 0 at y equal to the last knot (mflag /= 0, 631-636)
    v, fl = OA.bspline_eval3((x[3], ty[-1], z[3]), tx, ty, tz, coef, 4, 4, 4)
    assert v == 0.0


# ---------------------------------------------------------------- eigen / CG
def test_eig_max_against_numpy():
    rng = np.random.default_rng(11)
    for scale in (1e-8, 1.0, 1e6, 1e12):
        for _ in range(200):
            g = rng.normal(size=(3, 3)) * scale
            c = g.T @ g
            ref = np.linalg.eigvalsh(c).max()
            assert abs(OA.eig_sym3_max(c) - ref) <= 1e-13 * ref
    # degenerate and diagonal cases
    assert OA.eig_sym3_max(np.diag([3.0, 3.0, 3.0])) == 3.0
    assert OA.eig_sym3_max(np.zeros((3, 3))) == 0.0
    assert OA.eig_sym3_max(np.diag([1.0, 7.0, 2.0])) == 7.0


def test_cauchy_green_is_gtg():
    g = np.random.default_rng(2).normal(size=(3, 3))
    np.testing.assert_allclose(OA.cauchy_green(g), g.T @ g, rtol=1e-15, atol=1e-15)


# ---------------------------------------------------------------- grid
def test_periodic_grid_ghosts_shifts_and_delta():
    n = 8
    X, Y, Z, fl = F.abc_grid(n)
    o = Oracle(X, Y, Z, fl)
    (gx, gy, gz), gflag, bn = o.ghost_grid()
    h = 2 * np.pi / n
    L = 2 * np.pi
    xs = gx[:, 1, 1]
    np.testing.assert_allclose(xs[1:-1], X[:, 0, 0], rtol=0, atol=0)
    assert abs(xs[0] - (X[0, 0, 0] - h)) < 1e-13 and abs(xs[-1] - (X[-1, 0, 0] + h)) < 1e-13
    ps, bb = o.periodic_shift()
    ps = ps.reshape(3, 3, 3, 3)  # [k][j][i][d]
    assert abs(ps[1, 1, 0, 0] - L) < 1e-13 and abs(ps[1, 1, 2, 0] + L) < 1e-13
    assert abs(ps[2, 2, 2, 2] + L) < 1e-13 and ps[1, 1, 1].tolist() == [0, 0, 0]
    assert abs(bb[0] - (X.min() - h)) < 1e-13 and abs(bb[3] - (X.max() + h)) < 1e-13
    assert np.all(gflag == F.LCS_INTERNAL)
    for d in range(3):
        assert np.all(bn[d] == 0.0)
    o.close()


def test_wall_fakes_flags_and_normals_2d():
    X, Y, Z, fl = F.double_gyre_grid(16, 8)
    o = Oracle(X, Y, Z, fl)
    (gx, gy, gz), gflag, bn = o.ghost_grid()
    dx = 2.0 / 16
    # mirror fakes in x and y, 2-D offsets in z by the largest extent / max(n)
    assert abs(gx[0, 3, 1] - (X[0, 2, 0] - dx)) < 1e-14
    assert abs(gy[4, 0, 1] - (Y[3, 0, 0] - 1.0 / 8)) < 1e-14
    scaling = 2.0 / 16 * 15 / 16  # range of x (2 - dx) over max(n)
    assert abs(gz[4, 4, 0] - (Z[3, 3, 0] - scaling)) < 1e-14
    assert abs(gz[4, 4, 2] - (Z[3, 3, 0] + scaling)) < 1e-14
    # flags: WALL ghosts next to WALL nodes, LCS_2D above/below INTERNAL nodes
    assert gflag[0, 4, 1] == F.LCS_WALL and gflag[5, 4, 0] == F.LCS_2D and gflag[1, 4, 0] == F.LCS_WALL
    # outward unit normal on a straight wall, diagonal at a corner
    # (bcnorm is returned on the ghosted layout: node (i,j,k) is at [i+1, j+1, k+1])
    np.testing.assert_allclose([bn[0][1, 5, 1], bn[1][1, 5, 1], bn[2][1, 5, 1]], [-1, 0, 0], atol=1e-12)
    np.testing.assert_allclose([bn[0][9, 8, 1], bn[1][9, 8, 1]], [0, 1], atol=1e-12)
    c = np.array([bn[0][1, 1, 1], bn[1][1, 1, 1]])
    assert c[0] < 0 and c[1] < 0 and abs(np.hypot(*c) - 1) < 1e-12
    assert np.all([bn[d][6, 5, 1] == 0 for d in range(3)])
    o.close()


# ---------------------------------------------------------------- closed forms
@pytest.mark.parametrize("integ", [F.EULER, F.TRAPEZOIDAL, F.RK2, F.RK3, F.RK4])
@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC, F.TSE, F.TSE_LIMIT])
def test_uniform_flow_displacements(integ, interp):
    """u = c: every scheme moves by c dt per step; T-flowmaps are c T (FWD) and
    -c T (BKWD); forward particles wrap the periodic box."""
    X, Y, Z, fl = F.abc_grid(8)
    o = Oracle(X, Y, Z, fl)
    T, h, dt = 0.6, 0.3, 0.1
    hf = o.diagnostic_init(F.FTLE_FWD, 0, T, h, "f")
    hb = o.diagnostic_init(F.FTLE_BKWD, 0, T, h, "b")
    o.set_option("INTEGRATOR", integ)
    o.set_option("INTERPOLATOR", interp)
    c = (2.9, -1.7, 0.8)  # about half the box over the run: particles wrap
    one = np.ones(X.shape)
    for s in range(14):
        o.update(c[0] * one, c[1] * one, c[2] * one, s * dt)
        assert o.error == 0
    assert o.noutput(hf) == 4 and o.noutput(hb) == 4
    for d in range(3):
        assert np.abs(o.composed(hf)[d] - c[d] * T).max() < 1e-11
        assert np.abs(o.composed(hb)[d] + c[d] * T).max() < 1e-11
        # 14 updates = 13 CFD steps, 4 outputs every 3 steps: 1 step in the interval
        assert np.abs(o.flowmap(hf)[d] - c[d] * dt).max() < 1e-12
        assert np.abs(o.flowmap(hb)[d] + c[d] * dt).max() < 1e-12
    # CFL sub-cycles (lp_motion_m.f90:90-101): 1/delta = 1/dx, CFL 0.5
    dxg = 2 * np.pi / 8
    assert o.nsub == int(np.ceil(sum(abs(v) for v in c) / dxg * dt / 0.5))
    o.close()


def amplification(integ, z):
    """Per-subcycle growth factor of x' = lam x for each integrator as the
    reference codes them (lp_motion_m.f90:168-327)."""
    if integ in (F.EULER, F.TRAPEZOIDAL):
        return 1 + z
    if integ == F.RK2:
        return 1 + z + z * z / 2
    if integ == F.RK3:  # k3 = z (x + 2 k2): not classical RK3
        return 1 + z + 2 * z * z / 3 + z ** 3 / 6
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


@pytest.mark.parametrize("integ", [F.EULER, F.RK2, F.RK3, F.RK4])
@pytest.mark.parametrize("direction", [F.FTLE_FWD, F.FTLE_BKWD])
@pytest.mark.parametrize("interp", [F.LINEAR, F.TSE, F.TSE_LIMIT])
def test_linear_strain_flowmap_and_ftle(integ, direction, interp):
    """u = (lam x, -lam y, 0): LINEAR and TSE interpolation are exact (TSE's
    central-difference gradient is exact for linear data, and the value stays
    inside the limiter's 3x3x3 range), so every node whose trajectory stays
    inside the domain carries fm = (R^N - 1) x with R the scheme's
    amplification factor, and the displacement-gradient FTLE is
    log(max|R^N - 1|)/T (lcs_m.f90:137-176)."""
    lam, T, h, dt = 0.8, 0.5, 0.25, 0.05
    X, Y, Z = F.cell_grid((60, 60, 1), (-1, -1, 0), (1, 1, 1))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    fl[0], fl[-1], fl[:, 0], fl[:, -1] = F.LCS_SLIP, F.LCS_SLIP, F.LCS_SLIP, F.LCS_SLIP
    o = Oracle(X, Y, Z, fl)
    hd = o.diagnostic_init(direction, 0, T, h, "s")
    o.set_option("INTEGRATOR", integ)
    o.set_option("INTERPOLATOR", interp)
    u, v, w = lam * X, -lam * Y, 0 * X
    nsteps = 16
    for s in range(nsteps):
        o.update(u, v, w, s * dt)
        assert o.error == 0
    nsub = o.nsub
    sgn = 1 if direction == F.FTLE_FWD else -1
    z = sgn * lam * dt / nsub
    Rx, Ry = amplification(integ, z), amplification(integ, -z)
    Ntot = round(T / dt) * nsub
    cx, cy = Rx ** Ntot - 1, Ry ** Ntot - 1
    comp = o.composed(hd)
    ftle = o.ftle(hd)
    # the SLIP ghosts are zero-gradient copies, which is not linear: keep away
    # from the boundary layer that the repeated interpolation smears inward
    m = (np.abs(X) < 0.25) & (np.abs(Y) < 0.25)
    assert np.abs(comp[0][m] - cx * X[m]).max() < 1e-12
    assert np.abs(comp[1][m] - cy * Y[m]).max() < 1e-12
    assert np.abs(comp[2]).max() == 0.0
    ref = np.log(max(abs(cx), abs(cy))) / T
    assert np.abs(ftle[m] - ref).max() < 1e-10
    o.close()


@pytest.mark.parametrize("interp", [F.TSE, F.TSE_LIMIT])
def test_tse_equals_linear_on_general_linear_field(interp):
    """u = M x with a general (non-symmetric, all entries nonzero) M on a 3-D
    walled grid: trilinear and Taylor (TSE) interpolation both reproduce
    linear data exactly, so away from the non-linear ghost layer the TSE
    flowmaps equal the LINEAR ones.  Any dropped, transposed or mis-signed
    gradient term breaks the equality."""
    M = np.array([[0.3, -0.7, 0.4], [0.5, -0.2, -0.6], [-0.45, 0.35, 0.1]])
    X, Y, Z = F.cell_grid((18, 16, 14), (-1, -1, -1), (1, 1, 1))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    for sl in ((0,), (-1,), (slice(None), 0), (slice(None), -1), (slice(None), slice(None), 0),
               (slice(None), slice(None), -1)):
        fl[sl] = F.LCS_SLIP
    u = M[0, 0] * X + M[0, 1] * Y + M[0, 2] * Z
    v = M[1, 0] * X + M[1, 1] * Y + M[1, 2] * Z
    w = M[2, 0] * X + M[2, 1] * Y + M[2, 2] * Z
    res = {}
    for it in (F.LINEAR, interp):
        o = Oracle(X, Y, Z, fl)
        hf = o.diagnostic_init(F.FTLE_FWD, 0, 0.1, 0.05, "f")
        hb = o.diagnostic_init(F.FTLE_BKWD, 0, 0.1, 0.05, "b")
        o.set_option("INTEGRATOR", F.RK4)
        o.set_option("INTERPOLATOR", it)
        for s in range(4):
            o.update(u, v, w, s * 0.025)
            assert o.error == 0
        assert o.noutput(hf) == 1 and o.noutput(hb) == 1
        res[it] = (o.flowmap(hf), o.flowmap(hb), o.composed(hf), o.composed(hb))
        o.close()
    m = (np.abs(X) < 0.4) & (np.abs(Y) < 0.4) & (np.abs(Z) < 0.4)
    for a, b in zip(res[F.LINEAR], res[interp]):
        for d in range(3):
            assert np.abs(a[d][m]).max() > 1e-3
            assert np.abs(a[d][m] - b[d][m]).max() < 1e-13


# ---------------------------------------------------------------- invariants
def _invariants_closed_form(G, quirk=True):
    """compute_invariants (invariants_m.f90:15-198) of one velocity gradient
    G[c][d] = du_c/dx_d by matrix algebra: S = G^T + G, W = G^T - G,
    QS = -tr(SS)/8, RS = -tr(SSS)/24 (1/24 as a REAL(4) literal),
    Q = QS - tr(WW)/8, R = RS - (tr(WWS) [+ the reference's S13-for-S12
    term])/8, L2 from the LAPACK-style upper triangle of SS + WW with a zero
    (3,3) entry."""
    S, W = G.T + G, G.T - G
    c24 = float(np.float32(1.0) / np.float32(24.0))
    qs = -np.trace(S @ S) / 8
    rs = -c24 * np.trace(S @ S @ S)
    q = qs - np.trace(W @ W) / 8
    extra = W[1, 2] * W[2, 0] * (S[0, 2] - S[0, 1]) if quirk else 0.0
    r = rs - (np.trace(W @ W @ S) + extra) / 8
    d = q ** 3 + 6.75 * r * r
    cls = (1 if r > 0 else 2) if d > 0 else (3 if r > 0 else 4)
    A = S @ S + W @ W
    A[2, 2] = 0.0
    ev = np.linalg.eigvalsh(A, UPLO="U")
    with np.errstate(divide="ignore", invalid="ignore"):
        h = float(np.sqrt(-(q - qs) / qs))
    return [float(q > 0), float(rs > 0), float(cls), h, float(ev[1] < 0)]


@pytest.mark.parametrize("M", [
    # the reference's seventh R term (S13 for S12) moves D from class 2 to 1 here
    [[0.02, -0.33, -0.13], [0.84, -0.62, -0.7], [-0.98, 0.7, 0.98]],
    [[0.3, -0.7, 0.4], [0.5, -0.2, -0.6], [-0.45, 0.35, 0.1]],
    [[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 0.0]],      # solid rotation: vortex
    [[1.0, 0.0, 0.0], [0.0, -0.5, 0.0], [0.0, 0.0, -0.5]],     # axisymmetric strain
])
def test_invariants_of_linear_field(M):
    """u = M x on a walled grid: grad_sr1 is exact at nodes off the boundary,
    so every node there carries the closed-form invariants of G = M."""
    M = np.array(M)
    X, Y, Z = F.cell_grid((9, 8, 7), (-1, -1, -1), (1, 1, 1))
    fl = np.full(X.shape, F.LCS_WALL, dtype=np.int32)
    o = Oracle(X, Y, Z, fl)
    u = M[0, 0] * X + M[0, 1] * Y + M[0, 2] * Z
    v = M[1, 0] * X + M[1, 1] * Y + M[1, 2] * Z
    w = M[2, 0] * X + M[2, 1] * Y + M[2, 2] * Z
    got = o.invariants(u, v, w)
    want = _invariants_closed_form(M)
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    for q in (0, 1, 2, 4):
        assert np.all(got[q][inner] == want[q]), q
    if np.isfinite(want[3]):
        assert np.abs(got[3][inner] - want[3]).max() <= 1e-12 * max(1.0, abs(want[3]))
    o.close()
    if M[0, 0] == 0.02:
        assert _invariants_closed_form(M)[2] != _invariants_closed_form(M, quirk=False)[2]


def test_invariants_boundary_and_periodic_gradients():
    """Face neighbours: the periodic image on a periodic axis, zero beyond a
    wall (ugrid's ghosts); coordinates from the ghosted grid.  Checked by an
    independent numpy central difference of the same padded field."""
    X, Y, Z = F.cell_grid((10, 9, 8), (0, 0, 0), (2 * np.pi, 2 * np.pi, 2.0))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    fl[:, :, 0] = F.LCS_WALL
    fl[:, :, -1] = F.LCS_WALL
    o = Oracle(X, Y, Z, fl)
    vel = [np.sin(Y) + 0.3 * Z, np.cos(X) * (1 + Z), np.sin(X + Y) - 0.2 * Z * Z]
    got = o.invariants(*vel)
    (gx, gy, gz), _, _ = o.ghost_grid()
    want = [np.zeros(X.shape) for _ in range(5)]
    for i in range(10):
        for j in range(9):
            for k in range(8):
                G = np.zeros((3, 3))
                for c in range(3):
                    f = vel[c]
                    ip, im = (i + 1) % 10, (i - 1) % 10
                    jp, jm = (j + 1) % 9, (j - 1) % 9
                    G[c, 0] = (f[ip, j, k] - f[im, j, k]) / (gx[i + 2, j + 1, k + 1] - gx[i, j + 1, k + 1])
                    G[c, 1] = (f[i, jp, k] - f[i, jm, k]) / (gy[i + 1, j + 2, k + 1] - gy[i + 1, j, k + 1])
                    fp = f[i, j, k + 1] if k < 7 else 0.0
                    fm = f[i, j, k - 1] if k > 0 else 0.0
                    G[c, 2] = (fp - fm) / (gz[i + 1, j + 1, k + 2] - gz[i + 1, j + 1, k])
                r = _invariants_closed_form(G)
                for q in range(5):
                    want[q][i, j, k] = r[q]
    for q in (0, 1, 2, 4):
        assert np.array_equal(got[q], want[q]), q
    assert np.allclose(got[3], want[3], rtol=1e-12, atol=0)
    # no velocity slot: the zero field (H = 0/0)
    z = o.invariants()
    assert np.all(z[0] == 0) and np.all(z[1] == 0) and np.all(z[2] == 4) and np.all(z[4] == 0)
    assert np.all(np.isnan(z[3]))
    o.close()


def test_error_paths():
    X, Y, Z, fl = F.abc_grid(4)
    o = Oracle(X, Y, Z, fl)
    o.set_option("NOT_AN_OPTION", 1)
    assert o.error == 1
    o.close()
    o = Oracle(X, Y, Z, fl)
    o.set_param("CFL", 0.4)
    assert o.error == 0
    o.set_param("BOGUS", 1.0)
    assert o.error == 1
    o.close()
    # periodicity mismatch: INTERNAL on one face only
    fl2 = fl.copy()
    fl2[-1] = F.LCS_WALL
    o = Oracle(X, Y, Z, fl2)
    assert o.error == 1
    o.close()
    # wrong size on update
    o = Oracle(X, Y, Z, fl)
    o.n = np.array([4, 4, 5], dtype=np.int32)
    o.update(np.zeros((4, 4, 5)), np.zeros((4, 4, 5)), np.zeros((4, 4, 5)), 0.0)
    assert o.error == 1
    o.close()


# ------------------------------------- diagnostic grids: resolution != 0, AUX_GRID
def test_refined_and_coarsened_grid_coordinates():
    """new_sgrid_from_sgrid (sgrid_m.f90:445-761) on uniform grids: res > 0
    puts res points into every cell at fractions r/(res+1), including the
    periodic wrap cell after the last node but not after the last node of a
    wall-bounded axis; res < 0 keeps every (|res|+1)-th node; an axis with
    one node is kept; |res| >= n falls back to the flow grid."""
    h = 2 * np.pi / 8
    X, Y, Z, fl = F.abc_grid(8)
    o = Oracle(X, Y, Z, fl)
    h2 = o.diagnostic_init(F.FTLE_BKWD, 2, 1.0, 0.5, "r2")
    hm = o.diagnostic_init(F.FTLE_FWD, -1, 1.0, 0.5, "m1")
    hz = o.diagnostic_init(F.FTLE_FWD, 8, 1.0, 0.5, "z8")
    assert o.error == 0
    shape, (x, y, z) = o.lcs_grid(h2)
    assert shape == (24, 24, 24)
    assert np.abs(x[:, 3, 5] - (0.5 * h + np.arange(24) * h / 3)).max() < 1e-14
    assert np.abs(z[2, 3, :] - (0.5 * h + np.arange(24) * h / 3)).max() < 1e-14
    shape, (x, y, z) = o.lcs_grid(hm)
    assert shape == (4, 4, 4)
    assert np.array_equal(y[1, :, 2], Y[0, 1::2, 0])
    assert o.lcs_grid(hz)[0] == (8, 8, 8)
    o.close()
    # wall-bounded 2-D: n (res+1) - res points, z kept
    X, Y, Z, fl = F.double_gyre_grid(16, 8)
    o = Oracle(X, Y, Z, fl)
    hd = o.diagnostic_init(F.FTLE_FWD, 1, 1.0, 0.5, "w")
    shape, (x, y, z) = o.lcs_grid(hd)
    assert shape == (31, 15, 1)
    assert x[-1, 0, 0] == X[-1, 0, 0] and x[0, 0, 0] == X[0, 0, 0]
    assert np.abs(np.diff(x[:, 0, 0]) - 1.0 / 16).max() < 1e-14
    assert np.all(z == Z[0, 0, 0])
    o.close()


def test_aux_grid_coordinates():
    """create_aux_grid (lcs_m.f90:227-435) on a uniform grid: every node of
    set X0/X1 (Y0/Y1, Z0/Z1) is the node shifted by -/+ 0.2 h along x (y, z),
    the mean face-neighbour distance being h at every node; on a 2-D grid the
    out-of-plane sets are projected back onto the plane."""
    X, Y, Z, fl = F.abc_grid(6)
    h = 2 * np.pi / 6
    o = Oracle(X, Y, Z, fl)
    o.set_option("AUX_GRID", 1)
    hd = o.diagnostic_init(F.FTLE_FWD, 0, 1.0, 0.5, "a")
    assert o.error == 0
    base = (X, Y, Z)
    for a in range(6):
        shape, xyz = o.lcs_grid(hd, a + 1)
        ax, sgn = a % 3, (-1 if a < 3 else 1)
        for d in range(3):
            want = base[d] + (sgn * 0.2 * h if d == ax else 0.0)
            assert np.abs(xyz[d] - want).max() < 1e-14
    o.close()
    X, Y, Z, fl = F.double_gyre_grid(10, 5)
    o = Oracle(X, Y, Z, fl)
    o.set_option("AUX_GRID", 1)
    hd = o.diagnostic_init(F.FTLE_BKWD, 0, 1.0, 0.5, "a2")
    for a in (2, 5):
        _, xyz = o.lcs_grid(hd, a + 1)
        assert np.array_equal(xyz[2], Z) and np.array_equal(xyz[0], X)
    o.close()


@pytest.mark.parametrize("res", [0, 1, 2, -1])
@pytest.mark.parametrize("aux", [False, True])
@pytest.mark.parametrize("direction", [F.FTLE_FWD, F.FTLE_BKWD])
def test_diagnostic_grid_linear_flow_closed_form(res, aux, direction):
    """u = s (a x, b y, 0) with s = -1 forward (particles move inward) and
    s = +1 backward (departure points move inward), so no interpolation ever
    reads the wall ghosts: with LINEAR interpolation and RK2 every node of the
    diagnostic's grid (refined, coarsened or auxiliary) carries
    fm = (R^N - 1) x, R the amplification factor of the sub-cycle, N the
    number of sub-cycles over T (dt_factor = 1/(1+res) backward).  The FTLE
    is log(max|R^N - 1|)/T at every node with AUX_GRID and at every
    non-boundary node with grad_sr1 (whose boundary stencil reads ghosts)."""
    a, b, T, h, dt = 0.8, 0.3, 0.25, 0.125, 1.0 / 64
    # coarsening keeps nodes 2, 4, ..: an odd count keeps both faces INTERNAL
    # (one wall face only would be rejected as in init_sgrid)
    n = 33 if res < 0 else 32
    X, Y, Z = F.cell_grid((n, n, 1), (-1, -1, 0), (1, 1, 1))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    # one wall layer: on a 2-D grid every WALL node projects (the interior
    # test k > 1 .and. k < nk never holds, lp_tracking_m.f90:238), so an
    # inner wall layer would move particles
    fl[0], fl[-1], fl[:, 0], fl[:, -1] = F.LCS_WALL, F.LCS_WALL, F.LCS_WALL, F.LCS_WALL
    o = Oracle(X, Y, Z, fl)
    o.set_option("AUX_GRID", 1 if aux else 0)
    o.set_option("INTEGRATOR", F.RK2)
    o.set_option("INTERPOLATOR", F.LINEAR)
    hd = o.diagnostic_init(direction, res, T, h, "lin")
    s = -1.0 if direction == F.FTLE_FWD else 1.0
    for step in range(17):
        o.update(s * a * X, s * b * Y, 0 * X, step * dt)
        assert o.error == 0
    assert o.noutput(hd) == 2
    nsub = o.nsub
    # set_dt with dt_factor = 1/(1+res) backward (lcs_m.f90:40-44): the CFL
    # is (a|x| + b|y|)/h dt at the outermost nodes
    cfl = (a + b) * abs(X).max() * n / 2 * dt
    dtf = 1.0 / max(1 + res, 1) if direction == F.FTLE_BKWD else 1.0
    assert nsub == max(int(np.ceil(cfl / (0.5 * dtf))), 1)
    sdt = (dt if direction == F.FTLE_FWD else -dt) / nsub
    cx = amplification(F.RK2, s * a * sdt) ** (16 * nsub) - 1
    cy = amplification(F.RK2, s * b * sdt) ** (16 * nsub) - 1
    shape, (x, y, z) = o.lcs_grid(hd)
    comp = o.composed(hd)
    # a particle off its tracked flow node near a corner is projected along
    # that node's corner-leaning wall normal (set_lp_bc, lp_tracking_m.f90:
    # 237-252): grids other than the flow grid are checked inside two layers
    e = 2 if (res or aux) else 0
    cut = (slice(e, shape[0] - e), slice(e, shape[1] - e))
    assert np.abs(comp[0] - cx * x)[cut].max() < 1e-13
    assert np.abs(comp[1] - cy * y)[cut].max() < 1e-13
    ftle = o.ftle(hd)
    ref = np.log(max(abs(cx), abs(cy))) / T
    e = max(e, 0 if aux else 1)
    inner = ftle[e:shape[0] - e, e:shape[1] - e]
    assert np.abs(inner - ref).max() < 1e-10
    o.close()


def warped_grid(n, eps=0.05):
    """A non-rectilinear 2-D grid on about [-1, 1]^2 (nodes displaced by
    eps sin(pi y) in x and eps sin(pi x) in y), walls on the outer lines."""
    X, Y, Z = F.cell_grid((n, n, 1), (-1, -1, 0), (1, 1, 1))
    Xw = X + eps * np.sin(np.pi * Y)
    Yw = Y + eps * np.sin(np.pi * X)
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    fl[0], fl[-1], fl[:, 0], fl[:, -1] = F.LCS_WALL, F.LCS_WALL, F.LCS_WALL, F.LCS_WALL
    return Xw, Yw, Z, fl


@pytest.mark.parametrize("direction", [F.FTLE_FWD, F.FTLE_BKWD])
@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC, F.TSE])
def test_non_rectilinear_linear_flow_closed_form(direction, interp):
    """On a non-rectilinear grid every interpolation is TSE or TSE_LIMIT
    (lp_tracking_m.f90:319-328) with the face least-squares gradient
    (gradient_m.f90:114-295), and the FTLE gradient is least-squares too
    (lcs_m.f90:131-133).  Both are exact for a linear field whatever the
    node positions, so for u = s (a x, b y) (information moving inward, as
    in the rectilinear closed-form test) the composed map is (R^N - 1) x
    and the FTLE log(max|R^N - 1|)/T at nodes whose stencils stay clear of
    the walls' ghost values, which the repeated stencils carry inward by
    one node per sub-cycle."""
    a, b, T, h, dt = 0.8, 0.3, 0.125, 0.125, 1.0 / 64
    X, Y, Z, fl = warped_grid(48)
    o = Oracle(X, Y, Z, fl)
    o.set_option("INTEGRATOR", F.RK2)
    o.set_option("INTERPOLATOR", interp)
    hd = o.diagnostic_init(direction, 0, T, h, "w")
    s = -1.0 if direction == F.FTLE_FWD else 1.0
    for step in range(9):
        o.update(s * a * X, s * b * Y, 0 * X, step * dt)
        assert o.error == 0
    assert o.noutput(hd) == 1
    nsub = o.nsub
    sdt = (dt if direction == F.FTLE_FWD else -dt) / nsub
    cx = amplification(F.RK2, s * a * sdt) ** (8 * nsub) - 1
    cy = amplification(F.RK2, s * b * sdt) ** (8 * nsub) - 1
    comp = o.composed(hd)
    e = 8 * nsub + 3
    cut = (slice(e, 48 - e), slice(e, 48 - e))
    assert np.abs(comp[0] - cx * X)[cut].max() < 1e-12
    assert np.abs(comp[1] - cy * Y)[cut].max() < 1e-12
    ref = np.log(max(abs(cx), abs(cy))) / T
    assert np.abs(o.ftle(hd)[cut] - ref).max() < 1e-9
    o.close()


def test_non_rectilinear_interpolation_is_tse_limit():
    """A LINEAR run on a non-rectilinear grid gives exactly the TSE_LIMIT
    run (the interpolator is replaced), and differs from TSE on a curved
    field (the limiter acts)."""
    X, Y, Z, fl = warped_grid(24)
    out = {}
    for interp in (F.LINEAR, F.TSE_LIMIT, F.TSE):
        o = Oracle(X, Y, Z, fl)
        o.set_option("INTEGRATOR", F.RK2)
        o.set_option("INTERPOLATOR", interp)
        hd = o.diagnostic_init(F.FTLE_BKWD, 0, 0.2, 0.1, "w")
        for step in range(6):
            u, v, w = F.double_gyre_velocity(X + 1.0, Y + 0.5, step * 0.05)
            o.update(u, v, w, step * 0.05)
            assert o.error == 0
        out[interp] = o.flowmap(hd)
        o.close()
    for c in range(3):
        assert np.array_equal(out[F.LINEAR][c], out[F.TSE_LIMIT][c])
    assert max(np.abs(out[F.TSE][c] - out[F.TSE_LIMIT][c]).max() for c in range(2)) > 0
