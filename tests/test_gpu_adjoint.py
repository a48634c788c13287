"""GPU tests of the adjoint C-ABI pair (cfd2lcs_adjoint_forward_c /
cfd2lcs_adjoint_backward_c), which has no counterpart in the reference:

* its forward FTLE equals the library's own BKWD diagnostic with h = T on the
  same snapshots (bit-identical for linear interpolation, where composing a
  single slot is exact);
* the gradient of a scalar loss matches FP64 central finite differences in
  random directions over all snapshots (smooth cubic B-spline path to 1e-6
  relative, linear path to 1e-4, whose bracket switches are kinks);
* checkpointing every k steps changes nothing but memory;
* the tile-store and colour-launch scatters agree;
* the torch.autograd binding carries the gradient to flow parameters
  (double-gyre A, epsilon, omega), checked against finite differences of the
  parameters.
"""
import numpy as np
import pytest

import flows as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dp():
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("dp")
    yield lib
    lib.finalize()


def gyre_snapshots(X, Y, times, A=0.1, eps=0.1, om=2 * np.pi / 10):
    """(nsnap, n2, n1, n0) arrays of the double gyre on the grid (C order of
    (k, j, i) = Fortran order of the library's arrays)."""
    us, vs = [], []
    for t in times:
        u, v, _ = F.double_gyre_velocity(X, Y, t, A=A, eps=eps, omega=om)
        us.append(np.transpose(u, (2, 1, 0)))
        vs.append(np.transpose(v, (2, 1, 0)))
    us, vs = np.stack(us), np.stack(vs)
    return us, vs, np.zeros_like(us)


def abc_snapshots(X, Y, Z, times, amp=0.3):
    """ABC flow with a time-periodic modulation, so that every snapshot
    differs; (nsnap, n2, n1, n0)."""
    u0, v0, w0 = F.abc_velocity(X, Y, Z)
    us, vs, ws = [], [], []
    for t in times:
        m = 1.0 + amp * np.sin(2.0 * t)
        us.append(np.transpose(m * u0, (2, 1, 0)))
        vs.append(np.transpose(v0 + amp * np.cos(t) * np.sin(Z), (2, 1, 0)))
        ws.append(np.transpose(w0, (2, 1, 0)))
    return np.stack(us), np.stack(vs), np.stack(ws)


def setup(lib, case, interp, integ, n):
    lib.finalize()
    if case == "gyre":
        X, Y, Z, fl = F.double_gyre_grid(*n)
    else:
        X, Y, Z, fl = F.abc_grid(n)
    lib.init(X, Y, Z, fl)
    lib.set_option("INTEGRATOR", integ)
    lib.set_option("INTERPOLATOR", interp)
    return X, Y, Z, fl


def dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


def fwd(lib, us, vs, ws, times, ckpt=1):
    ftle = torch.empty(us.shape[1:], dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    assert lib.adjoint_forward(us.data_ptr(), vs.data_ptr(), ws.data_ptr(), us.shape[0], times, ckpt,
                               ftle.data_ptr()) == 0
    assert lib.error == 0
    return ftle


def bwd(lib, wts, us):
    gu, gv, gw = torch.empty_like(us), torch.empty_like(us), torch.empty_like(us)
    torch.cuda.synchronize()
    assert lib.adjoint_backward(wts.data_ptr(), gu.data_ptr(), gv.data_ptr(), gw.data_ptr()) == 0
    assert lib.error == 0
    return gu, gv, gw


CASES = [
    ("gyre", F.LINEAR, F.RK2, (32, 16)),
    ("gyre", F.CUBIC, F.RK4, (24, 12)),
    ("abc", F.LINEAR, F.RK3, 10),
    ("abc", F.CUBIC, F.RK4, 10),
]
# every tile width of the 3-D scatter: 4 (nearest), 6 (linear), 7 (quadratic
# B-spline), 8 (cubic)
CASES_ALL_TILES = CASES + [
    ("abc", F.QUADRATIC, F.RK4, 10),
    ("abc", F.NEAREST_NBR, F.RK2, 10),
    ("gyre", F.QUADRATIC, F.RK3, (24, 12)),
]


@pytest.mark.parametrize("case,interp,integ,n", CASES + [("abc", F.QUADRATIC, F.RK4, 10)])
def test_adjoint_forward_matches_bkwd_diagnostic(dp, case, interp, integ, n):
    import libcfd2lcs_amd as L
    X, Y, Z, fl = setup(dp, case, interp, integ, n)
    times = [0.05 * s for s in range(9)]
    snaps = gyre_snapshots(X, Y, times) if case == "gyre" else abc_snapshots(X, Y, Z, times)
    us, vs, ws = (dev(a) for a in snaps)
    f_adj = fwd(dp, us, vs, ws, times).cpu().numpy()
    T = times[-1] - times[0]
    setup(dp, case, interp, integ, n)
    hd = dp.diagnostic_init(L.FTLE_BKWD, 0, T, T, "b")
    for s, t in enumerate(times):
        u, v, w = (np.transpose(a[s], (2, 1, 0)) for a in snaps)
        dp.update(u, v, w, t)
    assert dp.error == 0 and dp.noutput(hd) == 1
    f_diag = np.transpose(dp.ftle(hd), (2, 1, 0))
    if interp == F.LINEAR:
        assert np.array_equal(f_adj, f_diag)
    else:
        assert np.abs(f_adj - f_diag).max() <= 1e-9


def _loss_and_grad(lib, us, vs, ws, times, wts, ckpt=1):
    f = fwd(lib, us, vs, ws, times, ckpt)
    L = float((wts * f).sum())
    return L, bwd(lib, wts, us)


@pytest.mark.parametrize("case,interp,integ,n", CASES)
def test_adjoint_gradient_matches_central_differences(dp, case, interp, integ, n):
    X, Y, Z, fl = setup(dp, case, interp, integ, n)
    times = [0.05 * s for s in range(7)]
    snaps = gyre_snapshots(X, Y, times) if case == "gyre" else abc_snapshots(X, Y, Z, times)
    us, vs, ws = (dev(a) for a in snaps)
    rng = np.random.default_rng(7)
    wts = dev(rng.standard_normal(us.shape[1:]))
    L0, (gu, gv, gw) = _loss_and_grad(dp, us, vs, ws, times, wts)
    nsub0 = dp.adjoint_stats()["subcycles"]
    smooth = interp == F.CUBIC
    h = 1e-6 if smooth else 1e-7
    tol = 1e-6 if smooth else 1e-4
    errs = []
    for trial in range(3):
        d = [dev(rng.standard_normal(us.shape)) for _ in range(3)]
        if case == "gyre":
            d[2].zero_()
        ad = float(sum((g * e).sum() for g, e in zip((gu, gv, gw), d)))
        Lp = float((wts * fwd(dp, us + h * d[0], vs + h * d[1], ws + h * d[2], times)).sum())
        assert dp.adjoint_stats()["subcycles"] == nsub0
        Lm = float((wts * fwd(dp, us - h * d[0], vs - h * d[1], ws - h * d[2], times)).sum())
        fd = (Lp - Lm) / (2 * h)
        errs.append(abs(ad - fd) / max(abs(fd), 1e-30))
        print(f"{case} interp {interp} integ {integ}: adjoint {ad:.10e} fd {fd:.10e} rel {errs[-1]:.2e}")
    assert max(errs) <= tol


@pytest.mark.parametrize("case,interp,integ,n", CASES_ALL_TILES)
def test_tile_store_and_colour_scatter_agree(dp, case, interp, integ, n):
    """The two scatter paths of the reverse sub-cycle, tile store + gather
    (the default) and in-place colour launches (used when the tile store
    would exceed its size cap), give the same gradient up to the order of
    summation."""
    import libcfd2lcs_amd as L
    X, Y, Z, fl = setup(dp, case, interp, integ, n)
    times = [0.05 * s for s in range(7)]
    snaps = gyre_snapshots(X, Y, times) if case == "gyre" else abc_snapshots(X, Y, Z, times)
    us, vs, ws = (dev(a) for a in snaps)
    wts = dev(np.random.default_rng(11).standard_normal(us.shape[1:]))
    out = []
    try:
        for colours in (False, True):
            dp.set_option("ADJOINT_COLOURS", L.LCS_TRUE if colours else L.LCS_FALSE)
            Lv, g = _loss_and_grad(dp, us, vs, ws, times, wts)
            out.append((Lv, torch.stack(g).cpu().numpy()))
    finally:
        dp.set_option("ADJOINT_COLOURS", L.LCS_FALSE)
    assert out[0][0] == out[1][0]
    scale = np.abs(out[0][1]).max()
    assert scale > 0
    assert np.abs(out[0][1] - out[1][1]).max() <= 1e-12 * scale


def test_checkpointing_changes_nothing(dp):
    X, Y, Z, fl = setup(dp, "abc", F.CUBIC, F.RK4, 10)
    times = [0.05 * s for s in range(8)]
    us, vs, ws = (dev(a) for a in abc_snapshots(X, Y, Z, times))
    wts = dev(np.random.default_rng(3).standard_normal(us.shape[1:]))
    ref = None
    for ck in (1, 3, 7):
        L, g = _loss_and_grad(dp, us, vs, ws, times, wts, ck)
        g = torch.stack(g).cpu().numpy()
        if ref is None:
            ref = (L, g)
        else:
            assert L == ref[0]
            assert np.abs(g - ref[1]).max() <= 1e-12 * np.abs(ref[1]).max()


def test_autograd_flow_parameters_double_gyre(dp):
    """dL/d(A, eps, omega) through torch autograd (snapshots built by torch
    from the parameters) against central differences of the parameters."""
    from libcfd2lcs_amd.autograd import BkwdFTLE
    X, Y, Z, fl = setup(dp, "gyre", F.LINEAR, F.RK2, (32, 16))
    times = [0.1 * s for s in range(6)]
    xs = dev(np.transpose(X, (2, 1, 0)))
    ys = dev(np.transpose(Y, (2, 1, 0)))

    def snaps(p):
        A, eps, om = p
        us, vs = [], []
        for t in times:
            a = eps * torch.sin(om * t)
            b = 1 - 2 * eps * torch.sin(om * t)
            f = a * xs ** 2 + b * xs
            dfdx = 2 * a * xs + b
            us.append(-np.pi * A * torch.sin(np.pi * f) * torch.cos(np.pi * ys))
            vs.append(np.pi * A * torch.cos(np.pi * f) * torch.sin(np.pi * ys) * dfdx)
        u, v = torch.stack(us), torch.stack(vs)
        return u, v, torch.zeros_like(u)

    target = BkwdFTLE.apply(dp, *snaps(torch.tensor([0.12, 0.15, 2 * np.pi / 10], dtype=torch.float64,
                                                    device="cuda:0")), times, 2).detach()

    def loss(p):
        return ((BkwdFTLE.apply(dp, *snaps(p), times, 2) - target) ** 2).sum()

    p = torch.tensor([0.1, 0.1, 2 * np.pi / 10], dtype=torch.float64, device="cuda:0", requires_grad=True)
    L = loss(p)
    L.backward()
    g = p.grad.cpu().numpy()
    with torch.no_grad():
        for a in range(3):
            h = 1e-7
            e = torch.zeros(3, dtype=torch.float64, device="cuda:0")
            e[a] = h
            fd = (float(loss(p + e)) - float(loss(p - e))) / (2 * h)
            print(f"param {a}: autograd {g[a]:.8e} fd {fd:.8e}")
            assert abs(g[a] - fd) <= 1e-4 * max(abs(fd), 1e-12) + 1e-10


def test_adjoint_against_fp64_tangent_linear_reference(dp):
    """<dL/du, du> from the GPU adjoint equals the directional derivative of
    the torch tangent-linear model (torch.func.jvp of an oracle-pinned
    restatement of the same discrete map), periodic 2-D, LINEAR, RK2."""
    import tlm_reference as TR
    from test_tlm_reference import periodic_case
    X, Y, Z, fl, times, us, vs = periodic_case()
    dp.finalize()
    dp.init(X, Y, Z, fl)
    dp.set_option("INTEGRATOR", F.RK2)
    dp.set_option("INTERPOLATOR", F.LINEAR)
    u = dev(np.stack([np.transpose(a, (2, 1, 0)) for a in us]))   # (S, 1, n1, n0)
    v = dev(np.stack([np.transpose(a, (2, 1, 0)) for a in vs]))
    w0 = torch.zeros_like(u)
    rng = np.random.default_rng(11)
    wts = dev(rng.standard_normal(u.shape[1:]))
    f_gpu = fwd(dp, u, v, w0, times)
    gu, gv, gw = bwd(dp, wts, u)
    x = torch.tensor(X[:, 0, 0], dtype=torch.float64)
    y = torch.tensor(Y[0, :, 0], dtype=torch.float64)
    ut, vt = u[:, 0].cpu(), v[:, 0].cpu()
    f_ref, subs = TR.bkwd_ftle_2d(ut, vt, times, x, y)
    assert np.abs(f_gpu[0].cpu().numpy() - f_ref.numpy()).max() <= 1e-10
    assert dp.adjoint_stats()["subcycles"] == sum(subs)
    wc = wts[0].cpu()

    def L(a, b):
        return (wc * TR.bkwd_ftle_2d(a, b, times, x, y, nsub=subs)[0]).sum()

    for trial in range(3):
        du = torch.tensor(rng.standard_normal(ut.shape), dtype=torch.float64)
        dv = torch.tensor(rng.standard_normal(vt.shape), dtype=torch.float64)
        _, tl = torch.func.jvp(L, (ut, vt), (du, dv))
        ad = float((gu[:, 0].cpu() * du).sum() + (gv[:, 0].cpu() * dv).sum())
        print(f"adjoint {ad:.12e} tangent-linear {float(tl):.12e}")
        assert abs(ad - float(tl)) <= 1e-9 * abs(float(tl))


def test_double_gyre_adjoint_full_size(dp):
    """The 2-D adjoint workload at full size: double gyre 256 x 128 with walls
    on [0,2] x [0,1], dt = 0.01, T = 10 (1000 steps), RK2 + LINEAR; loss =
    sum (FTLE - FTLE_target)^2 with the target from A = 0.12, eps = 0.15.
    Gradients with respect to (A, eps, omega) through torch autograd and with
    respect to all snapshots are checked against FP64 central differences
    (step 1e-6) in 8 seeded random unit directions, and the snapshot
    gradient against 4th-order central differences over a step sweep, to
    1e-5 at the step where the differences' own FP64 noise allows it."""
    from libcfd2lcs_amd.autograd import BkwdFTLE
    X, Y, Z, fl = setup(dp, "gyre", F.LINEAR, F.RK2, (256, 128))
    times = [0.01 * s for s in range(1001)]
    xs = dev(np.transpose(X, (2, 1, 0)))
    ys = dev(np.transpose(Y, (2, 1, 0)))
    tt = torch.tensor(times, dtype=torch.float64, device="cuda:0").view(-1, 1, 1, 1)

    def snaps(p):
        A, eps, om = p
        a = eps * torch.sin(om * tt)
        b = 1 - 2 * eps * torch.sin(om * tt)
        f = a * xs ** 2 + b * xs
        dfdx = 2 * a * xs + b
        u = -np.pi * A * torch.sin(np.pi * f) * torch.cos(np.pi * ys)
        v = np.pi * A * torch.cos(np.pi * f) * torch.sin(np.pi * ys) * dfdx
        return u, v, torch.zeros_like(u)

    with torch.no_grad():
        target = BkwdFTLE.apply(dp, *snaps(torch.tensor([0.12, 0.15, 2 * np.pi / 10], dtype=torch.float64,
                                                        device="cuda:0")), times, 50)

    def loss_of(u, v, w):
        return ((BkwdFTLE.apply(dp, u, v, w, times, 50) - target) ** 2).sum()

    p = torch.tensor([0.1, 0.1, 2 * np.pi / 10], dtype=torch.float64, device="cuda:0", requires_grad=True)
    u, v, w = snaps(p)
    u.retain_grad()
    v.retain_grad()
    L = loss_of(u, v, w)
    L.backward()
    gp = p.grad.cpu().numpy()
    gu, gv = u.grad.detach(), v.grad.detach()
    errs = []
    with torch.no_grad():
        for a in range(3):
            h = 1e-6
            e = torch.zeros(3, dtype=torch.float64, device="cuda:0")
            e[a] = h
            fd = (float(loss_of(*snaps(p + e))) - float(loss_of(*snaps(p - e)))) / (2 * h)
            errs.append(abs(gp[a] - fd) / abs(fd))
            print(f"param {a}: autograd {gp[a]:.10e} fd {fd:.10e} rel {errs[-1]:.2e}")
        u0, v0, w0 = (x.detach() for x in snaps(p))
        g = torch.Generator(device="cuda:0").manual_seed(4)
        sweep = {}
        for trial in range(8):
            du = torch.randn(u0.shape, generator=g, dtype=torch.float64, device="cuda:0")
            dv = torch.randn(v0.shape, generator=g, dtype=torch.float64, device="cuda:0")
            nrm = float(torch.sqrt((du ** 2).sum() + (dv ** 2).sum()))
            du, dv = du / nrm, dv / nrm
            ad = float((gu * du).sum() + (gv * dv).sum())
            Ls = lambda h: float(loss_of(u0 + h * du, v0 + h * dv, w0))
            # the stated check: central differences with step 1e-6
            fd2 = (Ls(1e-6) - Ls(-1e-6)) / 2e-6
            errs.append(abs(ad - fd2) / max(abs(fd2), 1e-300))
            # 4th-order central differences (h, 2h) over a step sweep: the
            # loss carries FP64 rounding of ~1e-12 and directional
            # derivatives are ~1e-3, so the difference quotient's noise is
            # ~1e-12 / h; the sweep shows it fall until truncation takes over
            for hh in (1e-6, 1e-5, 1e-4):
                fd4 = (8 * (Ls(hh) - Ls(-hh)) - (Ls(2 * hh) - Ls(-2 * hh))) / (12 * hh)
                sweep.setdefault(hh, []).append(abs(ad - fd4) / max(abs(fd4), 1e-300))
            print(f"direction {trial}: adjoint {ad:.12e} fd(1e-6) {fd2:.12e} rel {errs[-1]:.2e} "
                  + " ".join(f"fd4({hh:.0e}) rel {sweep[hh][-1]:.2e}" for hh in sweep))
    for hh, e in sweep.items():
        print(f"4th-order central differences, h = {hh:.0e}: median rel {np.median(e):.2e} max {max(e):.2e}")
    assert max(errs[:3]) <= 1e-5
    # step 1e-6 is noise-limited (~1e-12 / 2e-9): a loose bar there ...
    assert np.median(errs[3:]) <= 1e-4 and max(errs[3:]) <= 1e-3
    # ... and the snapshot gradient to 1e-5 at the step where the sweep's
    # noise (falling as 1/h) and truncation (growing with h) balance
    assert max(sweep[1e-5]) <= 1e-5


# ------------------------------------------------- FP32 adjoint, ABC parameters
@pytest.fixture(scope="module")
def sp():
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("sp")
    yield lib
    lib.finalize()


def smooth_noise(n, nsnap, sigma=0.05, width=4.0, seed=0):
    """nsnap x 3 seeded Gaussian fields on the periodic n^3 box, smoothed by
    a Gaussian of `width` cells (in Fourier space) and scaled to standard
    deviation sigma; (nsnap, 3, n, n, n) in C order of (k, j, i)."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n) * 2 * np.pi
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    filt = np.exp(-0.5 * width ** 2 * (KX ** 2 + KY ** 2 + KZ ** 2))
    out = np.empty((nsnap, 3, n, n, n))
    for s in range(nsnap):
        for c in range(3):
            f = np.fft.ifftn(np.fft.fftn(rng.standard_normal((n, n, n))) * filt).real
            out[s, c] = sigma * f / f.std()
    return out


def abc_param_snapshots(p, X, Y, Z, noise, dtype=torch.float64):
    """ABC(A, B, C) plus the fixed perturbations, built by torch from the
    parameter tensor p so that autograd reaches p; (nsnap, n, n, n) each."""
    xs, ys, zs = (dev(np.transpose(a, (2, 1, 0))) for a in (X, Y, Z))
    A, B, Cc = p[0], p[1], p[2]
    nz = dev(noise)
    u = A * torch.sin(zs) + Cc * torch.cos(ys) + nz[:, 0]
    v = B * torch.sin(xs) + A * torch.cos(zs) + nz[:, 1]
    w = Cc * torch.sin(ys) + B * torch.cos(xs) + nz[:, 2]
    return tuple(a.to(dtype) for a in (u, v, w))


def subdomain_weights(X, Y, Z):
    """loss = mean FTLE over [0, pi]^3: weights 1/count there, 0 elsewhere"""
    m = (X <= np.pi) & (Y <= np.pi) & (Z <= np.pi)
    return dev(np.transpose(m / m.sum(), (2, 1, 0)))


ABC3 = [np.sqrt(3.0), np.sqrt(2.0), 1.0]


def test_fp32_adjoint_against_fp64_central_differences(dp, sp):
    """The FP32 adjoint (float atomics, FP32 tile store and overflow fields)
    on the 3-D adjoint configuration at 16^3: perturbed ABC snapshots
    (sigma 0.05, smoothed over 4 cells), RK4 + cubic, checkpoints every 3
    steps, loss = mean FTLE over [0, pi]^3.  Its directional derivatives
    match FP64 central differences of the FP64 loss to 1e-4 relative (4e-6
    measured), and the whole snapshot gradient matches the FP64 adjoint to
    1e-4 of its largest entry (5e-6 measured)."""
    n, nsnap = 16, 9
    times = [0.02 * s for s in range(nsnap)]
    X, Y, Z, fl = F.abc_grid(n)
    noise = smooth_noise(n, nsnap)
    wts = subdomain_weights(X, Y, Z)
    p = torch.tensor(ABC3, dtype=torch.float64, device="cuda:0")
    grads = {}
    snaps = {}
    for lib, dt in ((dp, torch.float64), (sp, torch.float32)):
        lib.finalize()
        lib.init(X, Y, Z, fl)
        lib.set_option("INTEGRATOR", F.RK4)
        lib.set_option("INTERPOLATOR", F.CUBIC)
        us, vs, ws = abc_param_snapshots(p, X, Y, Z, noise, dt)
        snaps[dt] = (us, vs, ws)
        _, g = _loss_and_grad(lib, us, vs, ws, times, wts, ckpt=3)
        grads[dt] = torch.stack([a.double() for a in g])
    g64, g32 = grads[torch.float64], grads[torch.float32]
    field = float((g32 - g64).abs().max() / g64.abs().max())
    print(f"FP32 adjoint vs FP64 adjoint: max |diff| / max |g| = {field:.2e}")
    assert field <= 1e-4
    us, vs, ws = snaps[torch.float64]
    rng = np.random.default_rng(5)
    h = 1e-6
    for trial in range(3):
        d = [dev(rng.standard_normal(us.shape)) for _ in range(3)]
        ad32 = float(sum((g * e).sum() for g, e in zip(g32, d)))
        ad64 = float(sum((g * e).sum() for g, e in zip(g64, d)))
        Lp = float((wts * fwd(dp, us + h * d[0], vs + h * d[1], ws + h * d[2], times, 3)).sum())
        Lm = float((wts * fwd(dp, us - h * d[0], vs - h * d[1], ws - h * d[2], times, 3)).sum())
        fd = (Lp - Lm) / (2 * h)
        e32, e64 = abs(ad32 - fd) / abs(fd), abs(ad64 - fd) / abs(fd)
        print(f"direction {trial}: fd {fd:.10e} FP32 adjoint rel {e32:.2e} FP64 adjoint rel {e64:.2e}")
        assert e64 <= 1e-6 and e32 <= 1e-4


@pytest.mark.parametrize("variant", ["dp", "sp"])
def test_abc_parameter_gradient(dp, sp, variant):
    """dL/d(A, B, C) of the 3-D adjoint configuration (perturbed ABC, 16^3,
    RK4 + cubic, checkpoints every 4 steps, loss = mean FTLE over
    [0, pi]^3) through torch autograd and the adjoint pair, against FP64
    central differences of the parameters: 1e-5 relative with either
    library (FP64 ~1e-10, FP32 ~1e-7 measured)."""
    from libcfd2lcs_amd.autograd import BkwdFTLE
    lib = dp if variant == "dp" else sp
    dt = torch.float64 if variant == "dp" else torch.float32
    n, nsnap = 16, 9
    times = [0.02 * s for s in range(nsnap)]
    X, Y, Z, fl = F.abc_grid(n)
    noise = smooth_noise(n, nsnap, seed=1)
    wts = subdomain_weights(X, Y, Z)
    for L_ in (dp, sp):
        L_.finalize()
        L_.init(X, Y, Z, fl)
        L_.set_option("INTEGRATOR", F.RK4)
        L_.set_option("INTERPOLATOR", F.CUBIC)

    def loss(which, q, d):
        return (wts * BkwdFTLE.apply(which, *abc_param_snapshots(q, X, Y, Z, noise, d), times, 4)).sum()

    p = torch.tensor(ABC3, dtype=torch.float64, device="cuda:0", requires_grad=True)
    Lv = loss(lib, p, dt)
    Lv.backward()
    g = p.grad.cpu().numpy()
    tol = 1e-5
    with torch.no_grad():
        for a in range(3):
            h = 1e-6
            e = torch.zeros(3, dtype=torch.float64, device="cuda:0")
            e[a] = h
            fd = (float(loss(dp, p + e, torch.float64)) - float(loss(dp, p - e, torch.float64))) / (2 * h)
            rel = abs(g[a] - fd) / abs(fd)
            print(f"{variant} d/d{'ABC'[a]}: autograd {g[a]:.10e} fd {fd:.10e} rel {rel:.2e}")
            assert rel <= tol
