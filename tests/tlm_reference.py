"""FP64 tangent-linear reference for the adjoint tests (test infrastructure).

A torch restatement of the backward semi-Lagrangian FTLE of one window
(the BKWD diagnostic with h = T) for a 2-D periodic grid (all nodes
LCS_INTERNAL, one node along z), LINEAR interpolation and the RK2 scheme:

  per sub-cycle (flowmap_m.f90:17-104, lp_motion_m.f90:129-210):
    u0 = (1-c0) u_n + c0 u_np1 at the node          (first stage at x0)
    x1 = x0 + (dt/2) u0
    x_dep = x0 + dt * I[(1-ch) u_n + ch u_np1](x1)  (bracket of the node)
    fm_new(x0) = I[fm_old](x_dep) + x_dep - x0
  with periodic ghost copies and LINEAR brackets relative to the node
  (lp_tracking_m.f90:342-474); then FTLE = log(sqrt(lmax(G^T G))) / T with
  G the periodic central-difference gradient of fm (lcs_m.f90:66-193,
  gradient_m.f90:11-66) and its z column zero (2-D).

Written with differentiable torch operations, so torch.func.jvp of it is
the tangent-linear model of the same discrete map; the forward itself is
pinned to the C oracle in tests/test_tlm_reference.py.
"""
import math

import torch


def _ghost(f):
    """(..., n1, n0) -> (..., n1+2, n0+2), periodic copies."""
    f = torch.cat([f[..., -1:, :], f, f[..., :1, :]], dim=-2)
    return torch.cat([f[..., :, -1:], f, f[..., :, :1]], dim=-1)


def _bracket(xl, q, node):
    """LINEAR bracket of q relative to node index (ghosted) on line xl:
    lower index a and the clamped local coordinate t."""
    n = xl.shape[0] - 2
    a = torch.where(q > xl[node], node, node - 1).clamp(0, n)
    x0, x1 = xl[a], xl[a + 1]
    t = ((q - x0) / (x1 - x0)).clamp(0.0, 1.0)
    return a, t


def _interp(fg, xl, yl, qx, qy, ii, jj):
    """Bilinear value of ghosted fields fg (C, n1+2, n0+2) at (qx, qy) with
    brackets relative to nodes (ii, jj) (ghosted indices), in the term order
    of the trilinear sum whose z weight is exactly 1."""
    ax, tx = _bracket(xl, qx, ii)
    ay, ty = _bracket(yl, qy, jj)
    mx, my = 1.0 - tx, 1.0 - ty
    out = []
    for c in range(fg.shape[0]):
        f = fg[c]
        f4, f5 = f[ay, ax], f[ay, ax + 1]
        f6, f7 = f[ay + 1, ax], f[ay + 1, ax + 1]
        zero = torch.zeros_like(qx)
        # terms of f0..f3 carry the z weight 0
        v = zero + zero + zero + zero
        v = v + (mx * my) * f4
        v = v + (tx * my) * f5
        v = v + (mx * ty) * f6
        v = v + (tx * ty) * f7
        out.append(v)
    return out


def bkwd_ftle_2d(u, v, times, x, y, cfl=0.5, nsub=None):
    """u, v: (S, n1, n0) float64 snapshots (C order of (j, i)); times: S
    floats; x, y: interior coordinates (n0,), (n1,) of a uniform periodic
    grid.  Returns (ftle (n1, n0), nsub list)."""
    S, n1, n0 = u.shape
    hx, hy = float(x[1] - x[0]), float(y[1] - y[0])
    xl = torch.cat([x[:1] - hx, x, x[-1:] + hx])
    yl = torch.cat([y[:1] - hy, y, y[-1:] + hy])
    jj, ii = torch.meshgrid(torch.arange(1, n1 + 1), torch.arange(1, n0 + 1), indexing="ij")
    X, Y = xl[ii], yl[jj]
    fm = torch.zeros((2, n1 + 2, n0 + 2), dtype=torch.float64, device=u.device)
    subs = []
    for s in range(1, S):
        t_n, t_np1 = float(times[s - 1]), float(times[s])
        un = torch.stack([u[s - 1], v[s - 1]])
        unp1 = torch.stack([u[s], v[s]])
        if nsub is None:
            rate = float((unp1[0].detach().abs() / hx + unp1[1].detach().abs() / hy).max())
            ns = max(int(math.ceil(rate * (t_np1 - t_n) / cfl)), 1)
        else:
            ns = nsub[s - 1]
        subs.append(ns)
        ung, unp1g = _ghost(un), _ghost(unp1)
        dt = -((t_np1 - t_n) / ns)
        t = t_np1
        for _ in range(ns):
            c0 = min(max((t - t_n) / (t_np1 - t_n), 0.0), 1.0)
            ch = min(max((t + 0.5 * dt - t_n) / (t_np1 - t_n), 0.0), 1.0)
            w0 = 1.0 - c0
            u0 = [w0 * un[c] + c0 * unp1[c] for c in range(2)]
            x1 = X + (0.5 * dt) * u0[0]
            y1 = Y + (0.5 * dt) * u0[1]
            vh = (1.0 - ch) * ung + ch * unp1g
            uh = _interp(vh, xl, yl, x1, y1, ii, jj)
            xp = X + dt * uh[0]
            yp = Y + dt * uh[1]
            F = _interp(fm, xl, yl, xp, yp, ii, jj)
            fmi = torch.stack([F[0] + (xp - X), F[1] + (yp - Y)])
            fm = _ghost(fmi)
            t = t + dt
    fi = fm[:, 1:-1, 1:-1]
    T = float(times[-1] - times[0])
    gx = (torch.roll(fi, -1, dims=2) - torch.roll(fi, 1, dims=2)) / (2 * hx)
    gy = (torch.roll(fi, -1, dims=1) - torch.roll(fi, 1, dims=1)) / (2 * hy)
    # C = G^T G over the in-plane block (the z row and column are zero)
    cxx = gx[0] * gx[0] + gx[1] * gx[1]
    cxy = gx[0] * gy[0] + gx[1] * gy[1]
    cyy = gy[0] * gy[0] + gy[1] * gy[1]
    C = torch.stack([torch.stack([cxx, cxy], -1), torch.stack([cxy, cyy], -1)], -2)
    lam = torch.linalg.eigvalsh(C)[..., -1]
    return torch.log(torch.sqrt(lam)) / T, subs
