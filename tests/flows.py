"""Analytic grids and velocity fields for the configurations the reference's
example drivers use (examples/abc_flow_c.c:164-228, examples/double_gyre_c.c:164-234)
and for the closed-form known-answer tests.  Arrays are (ni, nj, nk), Fortran
order when flattened (i fastest), as the C API expects."""
import numpy as np

LCS_INTERNAL, LCS_INFLOW, LCS_OUTFLOW, LCS_WALL, LCS_SLIP, LCS_2D, LCS_MASK = range(7)
FTLE_FWD, FTLE_BKWD, LP_TRACER = 0, 1, 2
EULER, TRAPEZOIDAL, RK2, RK3, RK4 = 0, 1, 2, 3, 4
NEAREST_NBR, LINEAR, QUADRATIC, CUBIC, TSE, TSE_LIMIT = 0, 1, 2, 3, 5, 6


def cell_grid(n, lo, hi):
    """Cell-centred uniform grid: x_i = lo + (i + 1/2) (hi - lo) / n."""
    axes = []
    for d in range(3):
        h = (hi[d] - lo[d]) / max(n[d], 1)
        axes.append(lo[d] + 0.5 * h + np.arange(n[d]) * h)
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    return X, Y, Z


def abc_grid(n):
    L = 2.0 * np.pi
    X, Y, Z = cell_grid((n, n, n) if np.isscalar(n) else n, (0, 0, 0), (L, L, L))
    flag = np.full(X.shape, LCS_INTERNAL, dtype=np.int32)
    return X, Y, Z, flag


def abc_velocity(X, Y, Z, A=np.sqrt(3.0), B=np.sqrt(2.0), Cc=1.0):
    u = A * np.sin(Z) + Cc * np.cos(Y)
    v = B * np.sin(X) + A * np.cos(Z)
    w = Cc * np.sin(Y) + B * np.cos(X)
    return u, v, w


def double_gyre_grid(ni, nj, walls=True):
    """[0,2]x[0,1] cell-centred, one plane in z; LCS_WALL on the outermost
    i and j node lines."""
    X, Y, Z = cell_grid((ni, nj, 1), (0.0, 0.0, 0.0), (2.0, 1.0, 1.0))
    flag = np.full(X.shape, LCS_INTERNAL, dtype=np.int32)
    if walls:
        flag[0, :, :] = LCS_WALL
        flag[-1, :, :] = LCS_WALL
        flag[:, 0, :] = LCS_WALL
        flag[:, -1, :] = LCS_WALL
    return X, Y, Z, flag


def double_gyre_velocity(X, Y, t, A=0.1, eps=0.1, omega=2 * np.pi / 10):
    a = eps * np.sin(omega * t)
    b = 1.0 - 2.0 * eps * np.sin(omega * t)
    f = a * X ** 2 + b * X
    dfdx = 2.0 * a * X + b
    u = -np.pi * A * np.sin(np.pi * f) * np.cos(np.pi * Y)
    v = np.pi * A * np.cos(np.pi * f) * np.sin(np.pi * Y) * dfdx
    w = np.zeros_like(X)
    return u, v, w


class RandomFourierField:
    """Synthetic divergence-free turbulent velocity on the periodic
    cell-centred grid x_i = (i + 1/2) 2 pi / n: a sum of Fourier modes k
    (integer wavevectors, 0 < |k| <= kmax) with seeded uniform phases and
    seeded random directions projected perpendicular to k, amplitudes from
    E(k) ~ k^4 exp(-2 (k/4)^2) shared over each shell, the mode -k the
    complex conjugate of k (a real field), normalised to unit RMS of |u| at
    t = 0; each mode's phase advances at omega_k = |k|^(2/3).  Evaluated on
    the device with an inverse FFT (torch, FP64); at(t) returns (u, v, w) of
    shape (n, n, n) in C order of (k, j, i)."""

    def __init__(self, n, kmax=42, seed=0, device="cuda:0"):
        import torch
        self.torch = torch
        self.n = n
        dev = torch.device(device)
        rng = np.random.default_rng(seed)
        k1 = np.fft.fftfreq(n, d=1.0 / n)
        KZ, KY, KX = np.meshgrid(k1, k1, k1, indexing="ij")
        kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
        sel = (kk > 0) & (kk <= kmax)
        # one representative of each +-k pair: first nonzero component > 0
        half = (KX > 0) | ((KX == 0) & (KY > 0)) | ((KX == 0) & (KY == 0) & (KZ > 0))
        rep = sel & half
        idx = np.nonzero(rep)
        kx, ky, kz, km = KX[idx], KY[idx], KZ[idx], kk[idx]
        nm = kx.size
        d = rng.standard_normal((nm, 3))
        kv = np.stack([kx, ky, kz], 1) / km[:, None]
        d -= (d * kv).sum(1)[:, None] * kv
        d /= np.linalg.norm(d, axis=1)[:, None]
        phase = rng.uniform(0, 2 * np.pi, nm)
        shell = np.rint(km).astype(int)
        counts = np.bincount(shell, minlength=kmax + 2).astype(float)
        E = km ** 4 * np.exp(-2.0 * (km / 4.0) ** 2)
        amp = np.sqrt(E / counts[shell])
        h = 2 * np.pi / n
        # cell-centred offset e^{i k . h/2}
        coef = amp * np.exp(1j * (phase + 0.5 * h * (kx + ky + kz)))
        self.C = []
        for c in range(3):
            F = np.zeros((n, n, n), dtype=np.complex128)
            F[idx] = coef * d[:, c]
            mi = ((-idx[0]) % n, (-idx[1]) % n, (-idx[2]) % n)
            F[mi] = np.conj(coef * d[:, c])
            self.C.append(torch.tensor(F, device=dev))
        mi = ((-idx[0]) % n, (-idx[1]) % n, (-idx[2]) % n)
        # the phase of k advances by +omega t and that of its mirror -k by
        # -omega t, so the field stays real
        sgn = np.zeros((n, n, n))
        sgn[idx] = 1.0
        sgn[mi] = -1.0
        self.omega = torch.tensor(sgn * np.where(sel, kk, 0.0) ** (2.0 / 3.0), device=dev)
        self.scale = 1.0
        u = self.at(0.0)
        rms = float(torch.sqrt((u[0] ** 2 + u[1] ** 2 + u[2] ** 2).mean()))
        self.scale = 1.0 / rms

    def at(self, t):
        torch = self.torch
        rot = torch.exp(1j * self.omega * t)
        out = []
        for c in range(3):
            f = torch.fft.ifftn(self.C[c] * rot).real * (self.n ** 3) * self.scale
            out.append(f.contiguous())
        return out


def uniform_grid(n, h=1.0, lo=0.0):
    """Cell-centred grid with spacing h on every axis: x_i = lo + (i + 1/2) h
    (exact in binary for h = 1 and lo = 0); all flags INTERNAL."""
    n = (n, n, n) if np.isscalar(n) else n
    X, Y, Z = cell_grid(n, (lo, lo, lo), tuple(lo + h * m for m in n))
    return X, Y, Z, np.full(X.shape, LCS_INTERNAL, dtype=np.int32)


def channel_grid(ni, nj, nk=1, mask=True, walls=LCS_WALL):
    """[0,2]x[0,1](x[0,1]) cell-centred channel: LCS_INFLOW on the i = 1
    face, LCS_OUTFLOW on the i = ni face, `walls` on the j faces (and on
    the k faces in 3-D), and with mask an LCS_MASK block around (0.9, 0.55)."""
    X, Y, Z = cell_grid((ni, nj, nk), (0.0, 0.0, 0.0), (2.0, 1.0, 1.0))
    fl = np.full(X.shape, LCS_INTERNAL, dtype=np.int32)
    fl[:, 0, :] = walls
    fl[:, -1, :] = walls
    if nk > 1:
        fl[:, :, 0] = walls
        fl[:, :, -1] = walls
    fl[0, :, :] = LCS_INFLOW
    fl[-1, :, :] = LCS_OUTFLOW
    if mask:
        m = (np.abs(X - 0.9) < 0.14) & (np.abs(Y - 0.55) < 0.12)
        if nk > 1:
            m &= np.abs(Z - 0.5) < 0.3
        fl[m] = LCS_MASK
    return X, Y, Z, fl


def channel_velocity(X, Y, Z, t, U=0.3):
    """Mean flow U along x plus the double gyre; w = 0.1 sin(2 pi z) in 3-D."""
    u, v, w = double_gyre_velocity(X, Y, t)
    if X.shape[2] > 1:
        w = 0.1 * np.sin(2 * np.pi * Z) * np.cos(np.pi * Y)
    return u + U, v, w
