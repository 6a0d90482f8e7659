"""Thin torch.autograd harness over the engine's weighted PageRank and its adjoint (prh_weighted_pagerank /
prh_weighted_adjoint). torch is only the carrier of device buffers here: both directions run the engine's HIP
kernels on device pointers, with torch's current stream drained before each call (the engine works on its own
stream and drains it before returning).

    r = PageRankFunction.apply(weights, teleport, graph, options)

weights: FP32 (or FP32-representable FP64) edge weights in the graph's out-CSR edge order, on the GPU;
teleport: FP64 teleport vector on the GPU. Gradients come back in the inputs' dtypes (the adjoint computes them in
FP64, so FP64 inputs receive FP64 gradients).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _native
from ._native import prh_stats
from .engine import DeviceGraph, PagerankOptions

# torch ships its own HIP runtime under the soname the engine library links; if the engine was loaded first, the
# system runtime serves the process and torch reports no GPU. Import torch (or this module) before the engine.
if _native._engine is not None and not torch.cuda.is_available():
    raise RuntimeError("pagerank_amd.autograd: torch sees no GPU because the engine library was loaded before torch; "
                       "import torch before creating a pagerank_amd Context")


def _vp(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


class PageRankFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights: torch.Tensor, teleport: torch.Tensor, graph: DeviceGraph,
                options: PagerankOptions | None = None):
        o = options or PagerankOptions()
        if weights.device.type != "cuda" or teleport.device.type != "cuda":
            raise ValueError("weights and teleport must be GPU tensors")
        w32 = weights.detach().to(torch.float32).contiguous()
        s = teleport.detach().to(torch.float64).contiguous()
        if w32.numel() != graph.M or s.numel() != graph.N:
            raise ValueError("weights need one value per edge and teleport one per vertex")
        r = torch.empty(graph.N, dtype=torch.float64, device=s.device)
        torch.cuda.current_stream().synchronize()
        lib, h = graph.ctx._lib, graph.ctx.handle
        st = prh_stats()
        graph.ctx.check(lib.prh_weighted_pagerank(h, graph.handle, _vp(w32), _vp(s), C.byref(o.native()), _vp(r),
                                                  C.byref(st), 1))
        ctx.graph, ctx.options = graph, o
        ctx.dtypes = (weights.dtype, teleport.dtype)
        ctx.save_for_backward(w32, r)
        ctx.iterations = st.iterations
        return r

    @staticmethod
    def backward(ctx, grad_r: torch.Tensor):
        w32, r = ctx.saved_tensors
        graph, o = ctx.graph, ctx.options
        g = grad_r.detach().to(torch.float64).contiguous()
        gw = torch.empty(graph.M, dtype=torch.float64, device=r.device)
        gs = torch.empty(graph.N, dtype=torch.float64, device=r.device)
        torch.cuda.current_stream().synchronize()
        st = prh_stats()
        graph.ctx.check(graph.ctx._lib.prh_weighted_adjoint(graph.ctx.handle, graph.handle, _vp(w32), _vp(r), _vp(g),
                                                            C.byref(o.native()), None, _vp(gw), _vp(gs),
                                                            C.byref(st), 1))
        return gw.to(ctx.dtypes[0]), gs.to(ctx.dtypes[1]), None, None


def pagerank(weights: torch.Tensor, teleport: torch.Tensor, graph: DeviceGraph,
             options: PagerankOptions | None = None) -> torch.Tensor:
    return PageRankFunction.apply(weights, teleport, graph, options)
