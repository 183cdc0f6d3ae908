"""ZNSSD mismatch as a torch.autograd.Function backed by the HIP adjoint kernel.

``cost = znssd(P, coords, winsize)`` gives, per subset i, the ZNSSD mismatch
sum_p (fhat_p - ghat(W(P_i; p)))^2 of the frame resident in the context, and
its backward is the hand-written vector-Jacobian product through the bicubic
warp (staq_znssd_bwd_dev).  ``P`` is an (N, 6) float64 CUDA tensor
[u_x v_x u_y v_y u v]; gradients with respect to nodal displacements follow
from autograd when P is assembled from a nodal tensor (``P[:, 4:6] = U``).
"""
from __future__ import annotations

import torch

from ._lib import Context


class _Znssd(torch.autograd.Function):
    @staticmethod
    def forward(ctx, P, coords_cm, winsize, sctx: Context, wmax: int):
        if P.device.type != "cuda" or P.dtype != torch.float64:
            raise TypeError("P must be a float64 tensor on the GPU")
        N = P.shape[0]
        cost = torch.empty(N, dtype=torch.float64, device=P.device)
        status = torch.empty(N, dtype=torch.int32, device=P.device)
        Pc = P.contiguous()
        stream = torch.cuda.current_stream(P.device).cuda_stream
        sctx.check(sctx.L.staq_znssd_fwd_dev(sctx.h, N, wmax, Pc.data_ptr(), coords_cm.data_ptr(),
                                             winsize.data_ptr(), None, cost.data_ptr(),
                                             status.data_ptr(), stream), "staq_znssd_fwd_dev")
        ctx.save_for_backward(Pc, coords_cm, winsize)
        ctx.sctx = sctx
        ctx.wmax = wmax
        ctx.mark_non_differentiable(status)
        return cost, status

    @staticmethod
    def backward(ctx, gcost, gstatus):
        P, coords_cm, winsize = ctx.saved_tensors
        sctx = ctx.sctx
        N = P.shape[0]
        grad = torch.empty_like(P)
        g = gcost.contiguous().to(torch.float64)
        stream = torch.cuda.current_stream(P.device).cuda_stream
        sctx.check(sctx.L.staq_znssd_bwd_dev(sctx.h, N, ctx.wmax, P.data_ptr(),
                                             coords_cm.data_ptr(), winsize.data_ptr(), None,
                                             g.data_ptr(), grad.data_ptr(), stream),
                   "staq_znssd_bwd_dev")
        return grad, None, None, None, None


def znssd(P: torch.Tensor, coords: torch.Tensor, winsize: torch.Tensor, sctx: Context,
          wmax: int | None = None, coords_cm: torch.Tensor | None = None):
    """Per-subset ZNSSD mismatch (N,) and an out-of-image status (N,).

    coords: (N, 2) float64 CUDA tensor of node coordinates (1-based);
    winsize: (N,) int32 CUDA tensor.  ``wmax`` (the largest window) and
    ``coords_cm`` (coords as a (2, N) contiguous tensor) may be given to keep
    a repeated call free of a device-to-host read and of an allocation."""
    if coords_cm is None:
        coords_cm = coords.t().contiguous()  # N x 2 column-major = (2, N) row-major
    ws = winsize if winsize.dtype == torch.int32 and winsize.is_contiguous() else \
        winsize.to(torch.int32).contiguous()
    if wmax is None:
        wmax = int(ws.max().item())
    return _Znssd.apply(P, coords_cm, ws, sctx, int(wmax))
