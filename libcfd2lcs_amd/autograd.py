"""torch.autograd binding of the adjoint C-ABI pair.

``BkwdFTLE.apply(lib, u, v, w, times, checkpoint)`` returns the backward FTLE
field of one window (the BKWD diagnostic with h = T over the given velocity
snapshots) and, on ``backward()``, the gradient of a scalar loss with
respect to every snapshot through ``cfd2lcs_adjoint_backward_c``.  Snapshots
are device tensors of shape (nsnap, n2, n1, n0) -- C order of (k, j, i), i.e.
the Fortran order of the library -- in the library's precision.  Gradients
with respect to flow parameters follow from autograd's chain rule through
whatever torch code built the snapshots.
"""
import torch


class BkwdFTLE(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lib, u, v, w, times, checkpoint=1):
        if not (u.is_cuda and v.is_cuda and w.is_cuda):
            raise RuntimeError("BkwdFTLE: snapshots must be device tensors")
        u, v, w = (x.contiguous() for x in (u, v, w))
        nsnap = u.shape[0]
        ftle = torch.empty(u.shape[1:], dtype=torch.float64, device=u.device)
        torch.cuda.synchronize(u.device)
        rc = lib.adjoint_forward(u.data_ptr(), v.data_ptr(), w.data_ptr(), nsnap, list(times), checkpoint,
                                 ftle.data_ptr())
        if rc != 0 or lib.error:
            raise RuntimeError("cfd2lcs_adjoint_forward_c failed")
        ctx.lib = lib
        ctx.save_for_backward(u, v, w)
        return ftle

    @staticmethod
    def backward(ctx, gftle):
        u, v, w = ctx.saved_tensors
        g = gftle.detach().to(torch.float64).contiguous()
        gu, gv, gw = torch.empty_like(u), torch.empty_like(v), torch.empty_like(w)
        torch.cuda.synchronize(u.device)
        rc = ctx.lib.adjoint_backward(g.data_ptr(), gu.data_ptr(), gv.data_ptr(), gw.data_ptr())
        if rc != 0 or ctx.lib.error:
            raise RuntimeError("cfd2lcs_adjoint_backward_c failed")
        return None, gu, gv, gw, None, None
