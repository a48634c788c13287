"""Pins the oracle's branches that the smooth closed forms of test_oracle.py
never reach: tracking with the 0.99 hysteresis, the particle and flowmap
boundary conditions of every flag (WALL, SLIP, INFLOW, OUTFLOW, MASK, wall
corners), NEAREST_NBR, QUADRATIC, the TSE_LIMIT clamp, the first update and
INCOMPRESSIBLE.  Each expected value follows from the mathematics of a
constructed case (exact ties, symmetric normals, uniform flow against a
flat boundary, polynomial reproduction, a spike field), not from the oracle's
own formulas:

* track_lp2node (lp_tracking_m.f90:72-109): a node moves only if a
  neighbour is closer by the factor 0.99 in r^2, scanning the neighbours in
  NBR_OFFSET order (data_m.f90:123-152) with strict '<';
* set_lp_bc (lp_tracking_m.f90:185-277): WALL/SLIP on boundary nodes and
  MASK anywhere project onto the node's plane along bcnorm, the normalised
  least-squares gradient of min(flag, 1) (sgrid_m.f90:384-415); INFLOW and
  OUTFLOW boundary nodes stick;
* set_velocity_bc (sgrid_m.f90:810-866) zeroes MASK velocity and WALL
  ghosts; set_flowmap_bc (flowmap_m.f90:325-392) zeroes the flowmap on IB
  MASK / INFLOW / OUTFLOW nodes and WALL ghosts;
* NEAREST_NBR reads the tracked node (lp_tracking_m.f90:332-340); the
  interpolating B-splines of order 3 and 4 reproduce quadratics; TSE_LIMIT
  clamps to the 3x3x3 range (lp_tracking_m.f90:580-654);
* INCOMPRESSIBLE takes max(ftle, 0) (lcs_m.f90:183-185).
"""
import numpy as np
import pytest

import flows as F
import oracle_api as OA
from oracle_api import Oracle


@pytest.fixture(scope="module", autouse=True)
def _built():
    OA.lib()


def node_x(i):
    """coordinate of 1-based node i on F.uniform_grid (spacing 1, lo 0)"""
    return i - 0.5


# ---------------------------------------------------------------- tracking
@pytest.mark.parametrize("delta,start,want", [
    (0.5, 5, 5),      # exact tie: equal r^2 never moves
    (0.5, 6, 6),
    (0.5012, 5, 5),   # (0.4988/0.5012)^2 = 0.99044 > 0.99: hysteresis keeps the node
    (0.5013, 5, 6),   # (0.4987/0.5013)^2 = 0.98965 < 0.99: moves
    (0.4987, 6, 5),
    (0.4988, 6, 6),
])
def test_tracking_hysteresis_threshold(delta, start, want):
    """A point between nodes 5 and 6 moves to the other node only once r^2
    falls below 0.99 of the current node's: the switch lies at
    delta = 1 / (1 + sqrt(0.99)) = 0.501256 from either node."""
    X, Y, Z, fl = F.uniform_grid(10)
    o = Oracle(X, Y, Z, fl)
    q = (node_x(5) + delta, node_x(5), node_x(5))
    got, passes = o.track(q, (start, 5, 5))
    assert got == (want, 5, 5)
    assert passes == (1 if want == start else 2)
    o.close()


@pytest.mark.parametrize("qy,want", [(5.0, (6, 5, 5)), (4.0, (6, 5, 5)), (4.6, (6, 5, 5)), (5.2, (6, 6, 5))])
def test_tracking_neighbour_order_breaks_exact_ties(qy, want):
    """q = (x5 + 0.7, qy, z5) from node (5,5,5).  For qy = 5.0 (4.0) the face
    neighbour (6,5,5) and the edge neighbour (6,6,5) ((6,4,5)) are at exactly
    the same distance; the faces come before every edge in NBR_OFFSET and
    '<' is strict, so the face wins (a lexicographic (dk, dj, di) scan would
    reach (6,4,5) first).  Off the tie the nearer node wins."""
    X, Y, Z, fl = F.uniform_grid(10)
    o = Oracle(X, Y, Z, fl)
    got, _ = o.track((node_x(5) + 0.7, qy, node_x(5)), (5, 5, 5))
    assert got == want
    o.close()


def stretched_grid(n, seed, walls=True):
    rng = np.random.default_rng(seed)
    axes = [np.cumsum(rng.uniform(0.4, 1.6, m)) for m in n]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    if walls:
        for sl in ((0,), (-1,), (slice(None), 0), (slice(None), -1), (slice(None), slice(None), 0),
                   (slice(None), slice(None), -1)):
            fl[sl] = F.LCS_WALL
    return axes, X, Y, Z, fl


def test_tracking_fixed_point_on_stretched_grid():
    """Recursive tracking from a start up to two nodes off ends at a node
    that no neighbour beats by the factor 0.99 (the scan's fixed point), and
    that is the globally nearest node or within the hysteresis of it."""
    n = (11, 9, 7)
    axes, X, Y, Z, fl = stretched_grid(n, 3)
    o = Oracle(X, Y, Z, fl)
    (gx, gy, gz), _, _ = o.ghost_grid()
    rng = np.random.default_rng(4)
    for _ in range(400):
        q = np.array([rng.uniform(a[0], a[-1]) for a in axes])
        near = [int(np.argmin(np.abs(a - q[d]))) + 1 for d, a in enumerate(axes)]
        start = tuple(int(np.clip(near[d] + rng.integers(-2, 3), 1, n[d])) for d in range(3))
        got, passes = o.track(q, start)
        assert passes >= 1
        r2 = lambda i, j, k: (gx[i, j, k] - q[0]) ** 2 + (gy[i, j, k] - q[1]) ** 2 + (gz[i, j, k] - q[2]) ** 2
        here = r2(*got)
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for dk in (-1, 0, 1):
                    m = (np.clip(got[0] + di, 0, n[0] + 1), np.clip(got[1] + dj, 0, n[1] + 1),
                         np.clip(got[2] + dk, 0, n[2] + 1))
                    assert not (r2(*m) < 0.99 * here)
        best = min(r2(i, j, k) for i in range(1, n[0] + 1) for j in range(1, n[1] + 1) for k in range(1, n[2] + 1))
        assert here <= best / 0.99 ** 3
    o.close()


# ---------------------------------------------------------------- set_lp_bc
def bc_box():
    """8^3, spacing 1: INFLOW / OUTFLOW on the j faces, WALL on i = 1 and
    SLIP on i = 8 (assigned last, so they own the edges), z periodic; a MASK
    block on nodes 4..6 of every axis; one interior node flagged WALL and
    one flagged INFLOW."""
    X, Y, Z, fl = F.uniform_grid(8)
    fl[:, 0, :] = F.LCS_INFLOW
    fl[:, -1, :] = F.LCS_OUTFLOW
    fl[0, :, :] = F.LCS_WALL
    fl[-1, :, :] = F.LCS_SLIP
    fl[3:6, 3:6, 3:6] = F.LCS_MASK
    fl[5, 5, 1] = F.LCS_WALL      # node (6, 6, 2): interior
    fl[2, 6, 6] = F.LCS_INFLOW    # node (3, 7, 7): interior
    return X, Y, Z, fl


def _nx(node):
    return np.array([node_x(a) for a in node])


@pytest.mark.parametrize("node,normal", [
    ((1, 4, 4), (-1, 0, 0)),           # flat WALL face: outward normal
    ((8, 4, 4), (1, 0, 0)),            # flat SLIP face
    ((1, 1, 4), (-1, -1, 0)),          # corner of two boundary faces: the diagonal
    ((4, 5, 5), (1, 0, 0)),            # MASK block face: into the block
    ((5, 5, 4), (0, 0, 1)),
])
def test_lp_bc_projects_along_symmetric_normals(node, normal):
    """By the symmetry of min(flag, 1) about these nodes, bcnorm is the unit
    vector shown; a point beyond the node's plane is moved back onto it along
    bcnorm, a point on the inner side is left alone."""
    X, Y, Z, fl = bc_box()
    o = Oracle(X, Y, Z, fl)
    _, _, bn = o.ghost_grid()
    nrm = np.array(normal, dtype=float) / np.linalg.norm(normal)
    assert np.allclose([bn[d][node] for d in range(3)], nrm, atol=1e-14, rtol=0)
    x0 = _nx(node)
    tang = np.cross(nrm, [0.3, 0.7, 0.9])
    tang = 0.2 * tang / np.linalg.norm(tang)
    out = x0 + 0.35 * nrm + tang
    q, nd, stick = o.lp_bc(out, node)
    assert nd == node and not stick
    assert np.allclose(q, x0 + tang, atol=1e-14, rtol=0)
    inside = x0 - 0.35 * nrm + tang
    q, nd, stick = o.lp_bc(inside, node)
    assert np.array_equal(q, inside) and not stick
    o.close()


def test_lp_bc_stick_interior_flags_clamp_and_mask_core():
    X, Y, Z, fl = bc_box()
    o = Oracle(X, Y, Z, fl)
    p = np.array([2.2, -0.3, 4.1])
    # INFLOW / OUTFLOW boundary nodes stick to the node
    for node in ((4, 1, 4), (5, 8, 3)):
        q, nd, stick = o.lp_bc(p, node)
        assert stick and nd == node and np.array_equal(q, _nx(node))
    # the same flags on interior nodes do nothing, as does WALL there
    for node in ((3, 7, 7), (6, 6, 2)):
        q, nd, stick = o.lp_bc(p + 9.0, node)
        assert not stick and nd == node and np.array_equal(q, p + 9.0)
    # a ghost node is clamped to the boundary node first (then WALL applies)
    q, nd, stick = o.lp_bc(np.array([-0.2, 3.5, 3.5]), (0, 4, 4))
    assert nd == (1, 4, 4) and np.allclose(q, [0.5, 3.5, 3.5], atol=1e-15)
    # inside the MASK block bcnorm is 0: nothing moves
    q, nd, stick = o.lp_bc(p, (5, 5, 5))
    assert np.array_equal(q, p) and not stick
    o.close()


# ---------------------------------------------- uniform flow against boundaries
def run_box(flags_x, integ, interp, c=0.5, T=10.0, n=24, cfl=0.25, direction=F.FTLE_FWD):
    """n x 6 x 6 box of spacing 1, flags_x = (flag of node line i = 1, of
    i = n), y and z periodic, u = (c, 0, 0), dt = 1 (2 sub-cycles of 0.25
    cells at CFL 0.25), one output at T."""
    X, Y, Z, fl = F.uniform_grid((n, 6, 6))
    fl[0], fl[-1] = flags_x
    o = Oracle(X, Y, Z, fl)
    hd = o.diagnostic_init(direction, 0, T, T, "b")
    o.set_option("INTEGRATOR", integ)
    o.set_option("INTERPOLATOR", interp)
    o.set_param("CFL", cfl)
    one = np.ones(X.shape)
    for s in range(int(T) + 1):
        o.update(c * one, 0 * one, 0 * one, float(s))
        assert o.error == 0
    assert o.nsub == 2 and o.noutput(hd) == 1
    fm = o.composed(hd)
    o.close()
    return X, fm


@pytest.mark.parametrize("integ,interp", [(F.EULER, F.LINEAR), (F.RK2, F.LINEAR), (F.RK4, F.NEAREST_NBR),
                                          (F.RK3, F.NEAREST_NBR)])
@pytest.mark.parametrize("flag", [F.LCS_WALL, F.LCS_SLIP])
def test_forward_particles_stop_on_the_wall_node_plane(integ, interp, flag):
    """Uniform flow into a WALL (SLIP) boundary: a particle whose tracked node
    is the boundary node and that has passed its plane is projected back onto
    it, so forward particles end at min(x0 + c T, x_n) (x_n = 23.5)."""
    X, fm = run_box((flag, flag), integ, interp)
    x0 = X[:, 0, 0]
    want = np.minimum(x0 + 5.0, 23.5) - x0
    assert np.abs(fm[0] - want[:, None, None]).max() < 1e-12
    assert np.abs(fm[1]).max() == 0 and np.abs(fm[2]).max() == 0


@pytest.mark.parametrize("integ", [F.EULER, F.RK4])
def test_backward_flowmap_at_walls_nearest(integ):
    """Backward semi-Lagrangian step with NEAREST_NBR against a WALL: the
    upstream boundary node's departure point is projected back onto the node
    (displacement 0, fm = 0); every other node reads its own old flowmap and
    moves by -c dt, so fm = -c T exactly."""
    X, fm = run_box((F.LCS_WALL, F.LCS_WALL), integ, F.NEAREST_NBR, direction=F.FTLE_BKWD)
    assert np.abs(fm[0][0]).max() == 0
    assert np.abs(fm[0][1:] + 5.0).max() < 1e-12
    assert np.abs(fm[1]).max() == 0 and np.abs(fm[2]).max() == 0


def test_backward_flowmap_at_walls_linear():
    X, fm = run_box((F.LCS_WALL, F.LCS_WALL), F.RK2, F.LINEAR, direction=F.FTLE_BKWD)
    x0 = X[:, 0, 0]
    assert np.abs(fm[0][0]).max() == 0
    # a linear interpolation reads one cell upstream, so the wall's value
    # reaches at most one cell further per sub-cycle (20 sub-cycles)
    far = x0 > 0.5 + 20
    assert far.sum() == 3
    assert np.abs(fm[0][far] + 5.0).max() < 1e-12


@pytest.mark.parametrize("integ,interp", [(F.EULER, F.LINEAR), (F.RK2, F.LINEAR), (F.RK4, F.NEAREST_NBR)])
def test_inflow_outflow_particles_stick(integ, interp):
    """INFLOW at i = 1 and OUTFLOW at i = 24, u = c: particles on either face
    stick (dx = 0).  Particles at x0 = i - 1/2 move 1/4 cell per sub-cycle
    until the tracked node becomes 24, which happens after the sub-cycle
    that takes them past x23 + 0.50126 = 23.00126, i.e. at x = 23.25; from
    the next sub-cycle on they stick there, so dx = 23.75 - i for i >= 19
    and c T = 5 for i <= 18."""
    X, fm = run_box((F.LCS_INFLOW, F.LCS_OUTFLOW), integ, interp)
    i = np.arange(1, 25)
    want = np.where(i >= 19, 23.75 - i, 5.0)
    want[0] = 0.0
    want[-1] = 0.0
    assert np.abs(fm[0] - want[:, None, None]).max() < 1e-12
    assert np.abs(fm[1]).max() == 0 and np.abs(fm[2]).max() == 0


@pytest.mark.parametrize("interp", [F.NEAREST_NBR, F.LINEAR])
def test_inflow_outflow_backward_flowmap_is_zero_on_the_faces(interp):
    X, fm = run_box((F.LCS_INFLOW, F.LCS_OUTFLOW), F.RK2, interp, direction=F.FTLE_BKWD)
    x0 = X[:, 0, 0]
    assert np.abs(fm[0][0]).max() == 0 and np.abs(fm[0][-1]).max() == 0
    sel = slice(1, -1) if interp == F.NEAREST_NBR else (x0 > 0.5 + 20) & (x0 < 23.5)
    assert np.abs(fm[0][sel] + 5.0).max() < 1e-12


@pytest.mark.parametrize("interp", [F.LINEAR, F.NEAREST_NBR, F.CUBIC])
def test_mask_block_zero_velocity_and_flowmap(interp):
    """Periodic 12^3 box with a MASK block on nodes 4..8: the velocity there
    is zeroed, so forward particles inside the block's core (all 27
    neighbours MASK) never move, and the backward flowmap is 0 on every MASK
    node."""
    X, Y, Z, fl = F.uniform_grid(12)
    fl[3:8, 3:8, 3:8] = F.LCS_MASK
    o = Oracle(X, Y, Z, fl)
    hf = o.diagnostic_init(F.FTLE_FWD, 0, 4.0, 4.0, "f")
    hb = o.diagnostic_init(F.FTLE_BKWD, 0, 4.0, 4.0, "b")
    o.set_option("INTEGRATOR", F.RK4)
    o.set_option("INTERPOLATOR", interp)
    one = np.ones(X.shape)
    for s in range(5):
        o.update(0.3 * one, 0.2 * one, -0.1 * one, float(s))
        assert o.error == 0
    cf, cb = o.composed(hf), o.composed(hb)
    core = np.zeros(X.shape, bool)
    core[4:7, 4:7, 4:7] = True
    m = fl == F.LCS_MASK
    for d in range(3):
        if interp != F.CUBIC:
            assert np.abs(cf[d][core]).max() == 0
        assert np.abs(cb[d][m]).max() == 0
        assert np.abs(cf[d][~m]).max() > 0.1
    # the velocity the interpolants read is zero on MASK nodes
    u = o.interp(F.NEAREST_NBR, 0.3 * one, 0.2 * one, -0.1 * one, [[5.5, 5.5, 5.5], [1.5, 1.5, 1.5]],
                 [[6, 6, 6], [2, 2, 2]])
    assert np.array_equal(u[0], [0, 0, 0]) and np.allclose(u[1], [0.3, 0.2, -0.1], rtol=0, atol=0)
    o.close()


# ---------------------------------------------------------------- NEAREST_NBR
@pytest.mark.parametrize("integ", [F.EULER, F.TRAPEZOIDAL, F.RK2, F.RK3, F.RK4])
def test_nearest_backward_flowmap_is_minus_t_times_u(integ):
    """With NEAREST_NBR every stage of a backward step reads the origin node
    (the step never re-tracks) and the old flowmap is read at that node too,
    so in a steady flow fm(x) = -(t - t0) u(x) at every node, for every
    scheme, however non-uniform u is."""
    X, Y, Z, fl = F.abc_grid(10)
    o = Oracle(X, Y, Z, fl)
    hb = o.diagnostic_init(F.FTLE_BKWD, 0, 1.2, 0.6, "b")
    o.set_option("INTEGRATOR", integ)
    o.set_option("INTERPOLATOR", F.NEAREST_NBR)
    u = F.abc_velocity(X, Y, Z)
    for s in range(9):
        o.update(*u, s * 0.1)
        assert o.error == 0
    assert o.noutput(hb) == 1 and o.nsub > 1
    fm = o.flowmap(hb)  # 2 CFD steps after the output at t = 0.6
    for d in range(3):
        assert np.abs(fm[d] + 0.2 * u[d]).max() < 1e-13
    o.close()


# ---------------------------------------------------------- QUADRATIC / CUBIC
def quad_shear(interp, direction, a=1.0, b=0.05, T=0.2, dt=0.05):
    """u = (a y^2, b, 0) on a 2-D grid periodic in x with SLIP walls in y
    (8 x 64 x 1 cells on [0,1]^2).  Along a path y = y0 + b t, RK4 (Simpson)
    integrates a y^2 exactly, and both interpolating splines reproduce the
    quadratic velocity and the (quadratic in y) backward flowmap exactly
    away from the walls, where the zero-gradient ghosts are not quadratic."""
    X, Y, Z = F.cell_grid((8, 64, 1), (0, 0, 0), (1, 1, 1))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    fl[:, 0], fl[:, -1] = F.LCS_SLIP, F.LCS_SLIP
    o = Oracle(X, Y, Z, fl)
    hd = o.diagnostic_init(direction, 0, T, T, "q")
    o.set_option("INTEGRATOR", F.RK4)
    o.set_option("INTERPOLATOR", interp)
    u = (a * Y ** 2, b + 0 * Y, 0 * Y)
    for s in range(int(round(T / dt)) + 1):
        o.update(*u, s * dt)
        assert o.error == 0
    assert o.noutput(hd) == 1
    fm = o.composed(hd)
    o.close()
    y = Y[0, :, 0]
    if direction == F.FTLE_FWD:
        want = (a * ((y + b * T) ** 3 - y ** 3) / (3 * b), b * T)
    else:
        want = (-a * (y ** 3 - (y - b * T) ** 3) / (3 * b), -b * T)
    rows = slice(18, 46)
    return [np.abs(fm[d][:, rows, 0] - (want[d][None, rows] if d == 0 else want[d])).max() for d in range(2)]


@pytest.mark.parametrize("interp", [F.QUADRATIC, F.CUBIC])
@pytest.mark.parametrize("direction", [F.FTLE_FWD, F.FTLE_BKWD])
def test_splines_reproduce_a_quadratic_flow(interp, direction):
    ex, ey = quad_shear(interp, direction)
    assert ex < 1e-11 and ey < 1e-12


def test_linear_does_not_reproduce_the_quadratic_flow():
    ex, _ = quad_shear(F.LINEAR, F.FTLE_FWD)
    assert ex > 1e-7


# ---------------------------------------------------------------- TSE_LIMIT
def test_tse_limit_clamps_to_the_27_node_range():
    """A unit spike in u (and -1 in v) at node A = (5,5,5).  From node
    B = A + x^ the central gradient is -1/2 (+1/2 for v): TSE at x_B + 0.3
    gives -0.15 (+0.15), outside B's 3x3x3 range [0, 1] ([-1, 0]), so
    TSE_LIMIT clamps to 0; at x_B - 0.3 the value is inside and kept.  The
    range covers the corner neighbours: a ramp u = x with a spike of 10 at
    C - (1,1,1) leaves TSE at x_C + 1.4 (ramp value + 1.4) unclamped, and
    without the spike it is clamped to the largest neighbour value."""
    X, Y, Z, fl = F.uniform_grid(10)
    o = Oracle(X, Y, Z, fl)
    zero = np.zeros(X.shape)
    spike = zero.copy()
    spike[4, 4, 4] = 1.0
    B = (6, 5, 5)
    xb = _nx(B)
    q = [xb + [0.3, 0, 0], xb - [0.3, 0, 0]]
    tse = o.interp(F.TSE, spike, -spike, zero, q, [B, B])
    lim = o.interp(F.TSE_LIMIT, spike, -spike, zero, q, [B, B])
    assert np.allclose(tse[:, :2], [[-0.15, 0.15], [0.15, -0.15]], atol=1e-15)
    assert np.array_equal(lim[0, :2], [0.0, 0.0])
    assert np.allclose(lim[1, :2], [0.15, -0.15], atol=1e-15)
    C = (6, 6, 6)
    ramp = X.copy()
    ramp2 = ramp.copy()
    ramp2[4, 4, 4] = 10.0  # node C - (1,1,1)
    qc = [_nx(C) + [1.4, 0, 0]]
    a = o.interp(F.TSE_LIMIT, ramp2, zero, zero, qc, [C])
    b = o.interp(F.TSE_LIMIT, ramp, zero, zero, qc, [C])
    assert abs(a[0, 0] - (node_x(6) + 1.4)) < 1e-14
    assert abs(b[0, 0] - node_x(7)) < 1e-14
    o.close()


# ------------------------------------------------------- first call, INCOMPRESSIBLE
def test_first_update_moves_nothing():
    """The first update has t_n = t_np1 (cfd2lcs.f90:104-107): dt = 0, one
    sub-cycle, no integrations counted, and whatever the time blend ct is
    (0/0), all flowmaps stay exactly 0."""
    X, Y, Z, fl = F.double_gyre_grid(32, 16)
    o = Oracle(X, Y, Z, fl)
    hf = o.diagnostic_init(F.FTLE_FWD, 0, 1.0, 0.5, "f")
    hb = o.diagnostic_init(F.FTLE_BKWD, 0, 1.0, 0.5, "b")
    o.set_option("INTEGRATOR", F.RK4)
    o.update(*F.double_gyre_velocity(X, Y, 0.3), 0.3)
    assert o.error == 0 and o.nsub == 1
    assert o.integrations(1) == 0 and o.integrations(-1) == 0
    for h in (hf, hb):
        assert all(np.all(a == 0) for a in o.flowmap(h))
    o.close()


@pytest.mark.parametrize("lam,incomp", [(-0.8, 0), (-0.8, 1), (0.8, 1)])
def test_incompressible_clamps_negative_ftle(lam, incomp):
    """u = (lam x, lam y, 0) with RK2 + LINEAR (exact for linear data): fm =
    (R^N - 1) x per axis, FTLE = log|R^N - 1| / T.  A contraction (lam < 0)
    has |R^N - 1| < 1, a negative FTLE, which INCOMPRESSIBLE clamps to 0."""
    T, h, dt = 0.5, 0.25, 0.05
    X, Y, Z = F.cell_grid((60, 60, 1), (-1, -1, 0), (1, 1, 1))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    fl[0], fl[-1], fl[:, 0], fl[:, -1] = F.LCS_SLIP, F.LCS_SLIP, F.LCS_SLIP, F.LCS_SLIP
    o = Oracle(X, Y, Z, fl)
    hd = o.diagnostic_init(F.FTLE_FWD, 0, T, h, "s")
    o.set_option("INTEGRATOR", F.RK2)
    o.set_option("INCOMPRESSIBLE", incomp)
    for s in range(16):
        o.update(lam * X, lam * Y, 0 * X, s * dt)
        assert o.error == 0
    z = lam * dt / o.nsub
    R = (1 + z + z * z / 2) ** (round(T / dt) * o.nsub)
    ref = np.log(abs(R - 1)) / T
    want = max(ref, 0.0) if incomp else ref
    m = (np.abs(X) < 0.25) & (np.abs(Y) < 0.25)
    assert np.abs(o.ftle(hd)[m] - want).max() < 1e-10
    if lam < 0:
        assert ref < -0.5
    o.close()
