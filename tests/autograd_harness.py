"""torch.autograd wrappers around the library's adjoint entry points.

Gradient-check harness only: the product exposes the adjoints through the C
ABI (gg_spmv_bwd, gg_wpagerank_adjoint); these Functions let the tests drive
them with torch.autograd.grad / gradcheck.  Tensors live on the HIP device
torch calls "cuda"; the library runs on its own stream, so every call first
synchronises torch's stream, and every library call returns synchronised.
"""
import numpy as np
import torch

import graphgrind_amd as gg


def _sync():
    torch.cuda.synchronize()


class SpmvFn(torch.autograd.Function):
    """y = spmv(g, x) with the graph's FP32 weights held fixed."""

    @staticmethod
    def forward(ctx, x, g):
        _sync()
        y = torch.empty_like(x)
        gg.spmv(g, x.contiguous(), y)
        ctx.g = g
        ctx.save_for_backward(x)
        return y

    @staticmethod
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        _sync()
        gx = torch.empty_like(x)
        gg.spmv_bwd(ctx.g, x, gy.contiguous(), gx=gx, want_gw=False)
        return gx, None


class SpmvWeightFn(torch.autograd.Function):
    """y = spmv(g, x) as a function of the edge weights (CSR order); the
    forward loads w into the graph (weights are FP32 on the graph)."""

    @staticmethod
    def forward(ctx, w, x, g):
        g.set_weights(w.detach().float().cpu().numpy())
        _sync()
        y = torch.empty_like(x)
        gg.spmv(g, x.contiguous(), y)
        ctx.g = g
        ctx.save_for_backward(x)
        ctx.wdtype = w.dtype
        return y

    @staticmethod
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        _sync()
        gw = torch.empty(max(ctx.g.m, 1), dtype=x.dtype, device=x.device)
        gx = torch.empty_like(x)
        gg.spmv_bwd(ctx.g, x, gy.contiguous(), gx=gx, gw=gw)
        return gw[: ctx.g.m].to(ctx.wdtype), gx, None


class WPageRankFn(torch.autograd.Function):
    """r = wpagerank(g; w, v) with implicit-adjoint backward.  w (CSR order)
    is loaded into the graph as FP32; v is the teleport vector."""

    @staticmethod
    def forward(ctx, w, v, g, alpha, eps, max_iter, adj_eps):
        g.set_weights(w.detach().float().cpu().numpy())
        _sync()
        r = torch.empty_like(v)
        _, info = gg.wpagerank(g, np.float64 if v.dtype == torch.float64 else np.float32, alpha, eps, max_iter,
                               v=v.contiguous(), out=r)
        ctx.g, ctx.alpha, ctx.adj_eps, ctx.max_iter = g, alpha, adj_eps, max_iter
        ctx.iters = info.iters
        ctx.wdtype = w.dtype
        ctx.save_for_backward(r)
        return r

    @staticmethod
    def backward(ctx, u):
        (r,) = ctx.saved_tensors
        _sync()
        gw = torch.empty(max(ctx.g.m, 1), dtype=r.dtype, device=r.device)
        gv = torch.empty_like(r)
        gg.wpagerank_adjoint(ctx.g, r, u.contiguous(), ctx.alpha, ctx.adj_eps, ctx.max_iter, gw=gw, gv=gv)
        return gw[: ctx.g.m].to(ctx.wdtype), gv, None, None, None, None, None
