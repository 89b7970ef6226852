"""Weighted personalized PageRank as a differentiable torch op.

r = WeightedPageRank.apply(weights, personalization, ctx) runs the device
forward (the delta series r = sum_t delta_t, delta_0 = (1-d) p,
delta_{t+1} = d W^T delta_t with W the row-normalised edge weights) through the
C ABI; backward runs the hand-written adjoint, the same delta series on the
transpose graph, and returns dL/dweights (out-CSR order) and dL/dpersonalization.
No counterpart in the reference (its nearest app is the weighted personalized
PHP Ingress, php/php_ingress.h:15-120).  Tensors live on the context's GPU; there
is no CPU fallback.

torch and libsuminc must share one HIP runtime: import torch before the first
suminc_amd.Context is created (the library then binds to the runtime torch
loaded).  Loaded the other way round, torch does not see the GPU.
"""
from __future__ import annotations

import torch

from .context import Context


class WeightedPageRank(torch.autograd.Function):
    @staticmethod
    def forward(fctx, weights: torch.Tensor, personalization: torch.Tensor, sctx: Context):
        if not (weights.is_cuda and personalization.is_cuda):
            raise ValueError("WeightedPageRank needs device tensors (import torch before creating the "
                             "first suminc_amd.Context so that both share one HIP runtime)")
        w = weights.detach().to(torch.float32).contiguous()
        p = personalization.detach().to(torch.float32).contiguous()
        torch.cuda.current_stream().synchronize()  # inputs written by torch's stream
        sctx.wpr_set_weights(w)
        r = torch.empty_like(p)
        sctx.wpr_forward(p, out=r)  # synchronous
        fctx.sctx = sctx
        fctx.m = w.numel()
        # the library's backward uses the rank of the context's last forward: keep
        # this forward's inputs and its token to run it again if another forward
        # (or a weight change) came in between
        fctx.save_for_backward(w, p)
        fctx.token = sctx.wpr_token
        return r

    @staticmethod
    def backward(fctx, grad_r: torch.Tensor):
        g = grad_r.detach().to(torch.float32).contiguous()
        torch.cuda.current_stream().synchronize()
        sctx = fctx.sctx
        if sctx.wpr_token != fctx.token:
            w, p = fctx.saved_tensors
            sctx.wpr_set_weights(w)
            sctx.wpr_forward(p, out=torch.empty_like(p))
        gp = torch.empty_like(g)
        gw = torch.empty(fctx.m, dtype=torch.float32, device=g.device)
        sctx.wpr_backward(g, grad_p=gp, grad_w=gw)  # synchronous
        return gw, gp, None


def weighted_pagerank(weights: torch.Tensor, personalization: torch.Tensor, sctx: Context) -> torch.Tensor:
    return WeightedPageRank.apply(weights, personalization, sctx)
