"""MI355X-native PageRank engine (static, naive-dynamic, dynamic traversal, dynamic frontier, DF with pruning).

The compute path is libpagerank_hip.so (hand-written HIP for gfx950 behind the C ABI of include/pagerank_hip.h);
this package is its Python mirror plus host-side graph preparation.
"""
from .engine import (DYNAMIC_FRONTIER, DYNAMIC_TRAVERSAL, METHOD_NAMES, NAIVE_DYNAMIC, PRUNE_DYNAMIC_FRONTIER,
                     STATIC, BatchedResult, Context, DeviceGraph, PagerankError, PagerankOptions, PagerankResult,
                     Stream, Tuning, ppr_batched, weighted_adjoint, weighted_pagerank,
                     pagerankDynamicFrontierHip, pagerankDynamicTraversalHip, pagerankNaiveDynamicHip,
                     pagerankPruneDynamicFrontierHip, pagerankStaticHip, solve, solve_device)
from .graphs import Batch, HostGraph, TemporalReader, apply_batch, read_mtx_edges

__all__ = [
    "STATIC", "NAIVE_DYNAMIC", "DYNAMIC_TRAVERSAL", "DYNAMIC_FRONTIER", "PRUNE_DYNAMIC_FRONTIER", "METHOD_NAMES",
    "Context", "DeviceGraph", "PagerankError", "PagerankOptions", "PagerankResult", "Stream", "Tuning", "solve",
    "solve_device", "BatchedResult", "ppr_batched", "weighted_pagerank", "weighted_adjoint", "pagerankStaticHip", "pagerankNaiveDynamicHip", "pagerankDynamicTraversalHip",
    "pagerankDynamicFrontierHip", "pagerankPruneDynamicFrontierHip", "Batch", "HostGraph", "TemporalReader", "apply_batch", "read_mtx_edges",
]
