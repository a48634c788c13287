"""The allocation-free entry point cfd2lcs_adjoint_forward_noalloc_c,
called through the C ABI, against cfd2lcs_adjoint_forward_c on the same
snapshots: the same kernels in the same order give the same FTLE bit for
bit.  Two calls in a row, the second on a smaller grid with fewer
snapshots (the workspace keeps the first call's contents, which the second
must not see); a tape recorded before the allocation-free call still
drives the backward call afterwards; a grid too large for the workspace
fails with the error flag."""
import numpy as np
import pytest

import flows as F

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def snaps(n, nsnap, dtype, seed):
    X, Y, Z, fl = F.abc_grid(n)
    u0, v0, w0 = F.abc_velocity(X, Y, Z)
    rng = np.random.default_rng(seed)
    out = []
    for base in (u0, v0, w0):
        arr = np.stack([np.transpose(base * (1 + 0.1 * np.sin(0.7 * s)) + 0.05 * rng.standard_normal(base.shape),
                                     (2, 1, 0)) for s in range(nsnap)])
        out.append(torch.tensor(arr, dtype=dtype, device="cuda:0").contiguous())
    return (X, Y, Z, fl), out


def ftle_both(lib, n, nsnap, dtype, interp, seed):
    (X, Y, Z, fl), (u, v, w) = snaps(n, nsnap, dtype, seed)
    lib.finalize()
    lib.init(X, Y, Z, fl)
    lib.set_option("INTEGRATOR", F.RK4)
    lib.set_option("INTERPOLATOR", interp)
    times = [0.05 * s for s in range(nsnap)]
    fa = torch.empty((n, n, n), dtype=torch.float64, device="cuda:0")
    fb = torch.full((n, n, n), np.nan, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    assert lib.adjoint_forward_noalloc(u.data_ptr(), v.data_ptr(), w.data_ptr(), nsnap, times, fb.data_ptr()) == 0
    assert lib.error == 0
    assert lib.adjoint_forward(u.data_ptr(), v.data_ptr(), w.data_ptr(), nsnap, times, 2, fa.data_ptr()) == 0
    assert lib.error == 0
    torch.cuda.synchronize()
    return fa.cpu().numpy(), fb.cpu().numpy()


@pytest.mark.parametrize("variant", ["dp", "sp"])
@pytest.mark.parametrize("interp", [F.LINEAR, F.CUBIC])
def test_noalloc_forward_equals_adjoint_forward(variant, interp):
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs(variant)
    dt = torch.float64 if variant == "dp" else torch.float32
    for n, nsnap, seed in ((20, 9, 1), (12, 5, 2)):
        fa, fb = ftle_both(lib, n, nsnap, dt, interp, seed)
        assert np.isfinite(fa).all()
        assert np.array_equal(fa, fb), f"{variant} interp {interp} n {n}: max diff {np.abs(fa - fb).max():.3e}"
    lib.finalize()


def test_noalloc_call_leaves_the_tape_alone():
    """forward (tape) -> allocation-free forward on other snapshots ->
    backward gives the gradient of forward (tape) -> backward, up to the
    summation order of the backward's atomic overflow scatter."""
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("dp")
    n, nsnap = 12, 6
    (X, Y, Z, fl), (u, v, w) = snaps(n, nsnap, torch.float64, 3)
    _, (u2, v2, w2) = snaps(n, nsnap, torch.float64, 4)
    times = [0.05 * s for s in range(nsnap)]
    wts = torch.tensor(np.random.default_rng(5).standard_normal((n, n, n)), dtype=torch.float64, device="cuda:0")
    grads = []
    for extra in (False, True):
        lib.finalize()
        lib.init(X, Y, Z, fl)
        lib.set_option("INTEGRATOR", F.RK4)
        lib.set_option("INTERPOLATOR", F.CUBIC)
        f = torch.empty((n, n, n), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        assert lib.adjoint_forward(u.data_ptr(), v.data_ptr(), w.data_ptr(), nsnap, times, 2, f.data_ptr()) == 0
        if extra:
            f2 = torch.empty_like(f)
            torch.cuda.synchronize()
            assert lib.adjoint_forward_noalloc(u2.data_ptr(), v2.data_ptr(), w2.data_ptr(), nsnap, times,
                                               f2.data_ptr()) == 0
        g = [torch.empty_like(u) for _ in range(3)]
        torch.cuda.synchronize()
        assert lib.adjoint_backward(wts.data_ptr(), *(a.data_ptr() for a in g)) == 0
        assert lib.error == 0
        torch.cuda.synchronize()
        grads.append(torch.stack(g).cpu().numpy())
    scale = np.abs(grads[0]).max()
    assert scale > 0
    assert np.abs(grads[0] - grads[1]).max() <= 1e-12 * scale
    lib.finalize()


def test_noalloc_rejects_grids_beyond_the_workspace(capfd):
    import libcfd2lcs_amd as L
    lib = L.Cfd2lcs("sp")
    n, nsnap = 96, 2
    X, Y, Z, fl = F.abc_grid(n)
    lib.finalize()
    lib.init(X, Y, Z, fl)
    lib.set_option("INTERPOLATOR", F.CUBIC)
    u = torch.zeros((nsnap, n, n, n), dtype=torch.float32, device="cuda:0")
    f = torch.empty((n, n, n), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    assert lib.adjoint_forward_noalloc(u.data_ptr(), u.data_ptr(), u.data_ptr(), nsnap, [0.0, 0.02],
                                       f.data_ptr()) != 0
    assert lib.error != 0
    assert "MiB of workspace" in capfd.readouterr().out
    lib.finalize()
    assert lib.error == 0
