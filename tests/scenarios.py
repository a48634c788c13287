"""Shared scenario driver: run the same sequence of API calls on the oracle
and on the GPU library and collect what each exposes."""
import numpy as np

import flows as F
from oracle_api import Oracle


def make_case(name, n=None):
    if name == "abc":
        n = n or 16
        X, Y, Z, fl = F.abc_grid(n)
        vel = lambda t: F.abc_velocity(X, Y, Z)
    elif name == "gyre":
        ni, nj = n or (64, 32)
        X, Y, Z, fl = F.double_gyre_grid(ni, nj)
        vel = lambda t: F.double_gyre_velocity(X, Y, t)
    elif name == "gyre_slip":
        ni, nj = n or (64, 32)
        X, Y, Z, fl = F.double_gyre_grid(ni, nj, walls=False)
        fl[0, :, :] = F.LCS_SLIP
        fl[-1, :, :] = F.LCS_SLIP
        fl[:, 0, :] = F.LCS_SLIP
        fl[:, -1, :] = F.LCS_SLIP
        vel = lambda t: F.double_gyre_velocity(X, Y, t)
    elif name == "gyre_warp":
        # non-rectilinear: node lines bent by a few percent of the spacing
        ni, nj = n or (48, 24)
        X, Y, Z, fl = F.double_gyre_grid(ni, nj)
        X, Y = X + 0.02 * np.sin(2 * np.pi * Y), Y + 0.01 * np.sin(np.pi * X)
        vel = lambda t: F.double_gyre_velocity(X, Y, t)
    elif name == "abc_warp":
        # non-rectilinear and triply periodic: the warp has the box's period
        n = n or 12
        X, Y, Z, fl = F.abc_grid(n)
        X, Y, Z = X + 0.05 * np.sin(Y), Y + 0.05 * np.sin(Z), Z + 0.05 * np.sin(X)
        vel = lambda t: F.abc_velocity(X, Y, Z)
    elif name in ("channel", "channel3d"):
        # INFLOW / OUTFLOW faces, WALL (2-D) or SLIP (3-D) side faces, an
        # interior MASK block; mean flow plus the double gyre
        if name == "channel":
            ni, nj = n or (64, 32)
            X, Y, Z, fl = F.channel_grid(ni, nj)
        else:
            ni, nj, nk = n or (24, 16, 12)
            X, Y, Z, fl = F.channel_grid(ni, nj, nk, walls=F.LCS_SLIP)
        vel = lambda t: F.channel_velocity(X, Y, Z, t)
    elif name in ("wall_box", "io_box"):
        # uniform flow u = (1/2, 0, 0) along a box of spacing 1 with WALL
        # (INFLOW / OUTFLOW) node lines at i = 1 and i = n, y and z periodic
        X, Y, Z, fl = F.uniform_grid(n or (24, 6, 6))
        fl[0], fl[-1] = (F.LCS_WALL, F.LCS_WALL) if name == "wall_box" else (F.LCS_INFLOW, F.LCS_OUTFLOW)
        vel = lambda t: (0.5 + 0 * X, 0 * X, 0 * X)
    elif name == "abc_slip_y":
        # ABC box periodic in x and z, SLIP walls on the j faces: the ghost z
        # coordinates take the reference's periodic_shift(i,k,k,3) reading
        # (comms_m.f90:750), which here differs from periodic_shift(i,j,k,3)
        X, Y, Z, fl = F.abc_grid(n or (12, 10, 8))
        fl[:, 0, :] = F.LCS_SLIP
        fl[:, -1, :] = F.LCS_SLIP
        vel = lambda t: F.abc_velocity(X, Y, Z)
    elif name == "mask_box":
        # periodic box with a MASK block on nodes 4..8; a sheared mean flow,
        # so that the FTLE outside the block is well above rounding
        X, Y, Z, fl = F.uniform_grid(n or 12)
        fl[3:8, 3:8, 3:8] = F.LCS_MASK
        L = X.shape[0]
        vel = lambda t: (0.3 + 0.2 * np.sin(2 * np.pi * Y / L), 0.2 + 0.1 * np.sin(2 * np.pi * Z / L),
                         -0.1 + 0.1 * np.sin(2 * np.pi * X / L))
    elif name == "quad_shear":
        # u = (y^2, 0.05, 0): periodic in x, SLIP walls in y, 2-D
        X, Y, Z = F.cell_grid(n or (8, 64, 1), (0, 0, 0), (1, 1, 1))
        fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
        fl[:, 0], fl[:, -1] = F.LCS_SLIP, F.LCS_SLIP
        vel = lambda t: (Y ** 2, 0.05 + 0 * Y, 0 * Y)
    else:
        raise ValueError(name)
    return X, Y, Z, fl, vel


def run_both(lib, case, n=None, integrator=F.RK2, interpolator=F.LINEAR, T=1.0, h=0.25, dt=0.05,
             nsteps=30, diags=(F.FTLE_FWD, F.FTLE_BKWD), cfl=None, every_step=None, resolution=0,
             aux=False, stats=None, options=()):
    """Returns a list of per-step records {step, fm: {handle: (gpu, oracle)}, ftle ...}.
    resolution / aux: the diagnostics' grid (new_sgrid_from_sgrid) and
    AUX_GRID, both read when the diagnostics are created; stats, when a
    list, receives the integration counters of both sides at the end;
    options: further (name, value) set_option calls."""
    X, Y, Z, fl, vel = make_case(case, n)
    o = Oracle(X, Y, Z, fl)
    lib.finalize()
    lib.init(X, Y, Z, fl)
    if aux:
        o.set_option("AUX_GRID", 1)
        lib.set_option("AUX_GRID", 1)
    handles = []
    for d in diags:
        label = "fwd" if d == F.FTLE_FWD else "bkwd"
        ho = o.diagnostic_init(d, resolution, T, h, label)
        hg = lib.diagnostic_init(d, resolution, T, h, label)
        assert ho == hg
        assert lib.error == 0 and o.error == 0
        shape_g, xyz_g = lib.lcs_grid(hg)
        shape_o, xyz_o = o.lcs_grid(ho)
        assert shape_g == shape_o
        for c in range(3):
            assert np.array_equal(xyz_g[c], xyz_o[c]) or lib.dtype != np.float64
        handles.append(hg)
    for name, val in (("INTEGRATOR", integrator), ("INTERPOLATOR", interpolator)) + tuple(options):
        o.set_option(name, val)
        lib.set_option(name, val)
    if cfl is not None:
        o.set_param("CFL", cfl)
        lib.set_param("CFL", cfl)
    recs = []
    prev = {hd: 0 for hd in handles}
    for s in range(nsteps):
        t = s * dt
        u, v, w = vel(t)
        o.update(u, v, w, t)
        lib.update(u, v, w, t)
        assert o.error == 0, f"oracle error at step {s}"
        assert lib.error == 0, f"library error at step {s}"
        rec = {"step": s, "t": t, "fm": {}, "out": {}}
        if every_step is None or s % every_step == 0 or s == nsteps - 1:
            for hd in handles:
                rec["fm"][hd] = (lib.flowmap(hd), o.flowmap(hd))
        for hd in handles:
            if lib.noutput(hd) != o.noutput(hd):
                raise AssertionError(f"output count differs at step {s}: {lib.noutput(hd)} vs {o.noutput(hd)}")
        for hd in handles:
            if o.noutput(hd) > prev[hd]:
                rec["out"][hd] = (lib.composed(hd), o.composed(hd), lib.ftle(hd), o.ftle(hd))
                prev[hd] = o.noutput(hd)
        recs.append(rec)
    if stats is not None:
        st = lib.stats()
        stats.append(((st["integrations_fwd"], st["integrations_bkwd"]),
                      (o.integrations(1), o.integrations(-1))))
    o.close()
    return recs, handles


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)
