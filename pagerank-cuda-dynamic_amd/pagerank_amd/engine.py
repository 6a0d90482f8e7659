"""Python mirror of the reference's GPU PageRank interface over the engine's C ABI (include/pagerank_hip.h).

Names, option fields and result fields follow inc/pagerank.hxx (PagerankOptions :24-55, PagerankResult :64-113)
and the five entry points of inc/pagerankCuda.hxx (:864-988). The reference's DiGraph pair (x, xt) / (y, yt)
becomes a HostGraph (both CSRs); uploads happen once per graph and stay resident on the device.
Every failure raises PagerankError with the library's message; nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native
from ._native import f64p, prh_options, prh_stats, prh_tuning, u32p, u8p
from .graphs import Batch, HostGraph

STATIC, NAIVE_DYNAMIC, DYNAMIC_TRAVERSAL, DYNAMIC_FRONTIER, PRUNE_DYNAMIC_FRONTIER = range(5)
METHOD_NAMES = {
    STATIC: "pagerankStaticHip",
    NAIVE_DYNAMIC: "pagerankNaiveDynamicHip",
    DYNAMIC_TRAVERSAL: "pagerankDynamicTraversalHip",
    DYNAMIC_FRONTIER: "pagerankDynamicFrontierHip",
    PRUNE_DYNAMIC_FRONTIER: "pagerankPruneDynamicFrontierHip",
}
STATUS = {0: "success", 1: "invalid argument", 2: "invalid graph", 3: "out of memory", 4: "device error",
          5: "unsupported"}


class PagerankError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


@dataclass
class PagerankOptions:
    """PagerankOptions<double> (pagerank.hxx:24-55), with the reference's defaults."""
    repeat: int = 1
    tolerance: float = 1e-10
    frontierTolerance: float = 1e-6
    pruneTolerance: float = 1e-6
    damping: float = 0.85
    maxIterations: int = 500

    def native(self) -> prh_options:
        return prh_options(self.repeat, self.tolerance, self.frontierTolerance, self.pruneTolerance, self.damping,
                           self.maxIterations)


@dataclass
class PagerankResult:
    """PagerankResult<double> (pagerank.hxx:64-113); times in milliseconds averaged over repeat. preparationTime is
    the one-time build of the graph's layout and launch plan that the first solve using them reports (included in
    initializationTime and time)."""
    ranks: np.ndarray = field(default_factory=lambda: np.zeros(0))
    iterations: int = 0
    time: float = 0.0
    initializationTime: float = 0.0
    markingTime: float = 0.0
    computationTime: float = 0.0
    preparationTime: float = 0.0


@dataclass
class Tuning:
    switch_degree: int = 64
    huge_degree: int = 16384
    low_width: int = 0
    check_interval: int = 8
    source_slices: int = 0
    block_edges: int = 2048
    extension_direction: int = 1
    blocking: int = 0
    layout: int = 0
    window: int = 0


def _tuning_of(n) -> Tuning:
    return Tuning(n.switch_degree, n.huge_degree, n.low_width, n.check_interval, n.source_slices, n.block_edges,
                  n.extension_direction, n.blocking, n.layout, n.window)


class Context:
    """A device, its stream and its reusable workspace (prh_context)."""

    def __init__(self, device: int = 0):
        self._lib = _native.engine()
        h = C.c_void_p()
        st = self._lib.prh_context_create(device, C.byref(h))
        if st != 0:
            raise PagerankError(st, f"cannot create a context on device {device}")
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self._lib.prh_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def check(self, status: int):
        if status != 0:
            msg = self._lib.prh_context_last_error(self._h)
            raise PagerankError(status, msg.decode() if msg else "")

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return int(self._lib.prh_context_stream(self._h) or 0)

    def set_memory_limit(self, nbytes: int):
        self.check(self._lib.prh_context_set_memory_limit(self._h, nbytes))

    def memory_used(self) -> int:
        return int(self._lib.prh_context_memory_used(self._h))

    def set_tuning(self, t: Tuning):
        self.check(self._lib.prh_context_set_tuning(
            self._h, C.byref(prh_tuning(t.switch_degree, t.huge_degree, t.low_width, t.check_interval,
                                        t.source_slices, t.block_edges, t.extension_direction, t.blocking,
                                        t.layout, t.window))))

    def get_tuning(self) -> Tuning:
        n = prh_tuning()
        self.check(self._lib.prh_context_get_tuning(self._h, C.byref(n)))
        return _tuning_of(n)

    def autotune(self, yg: "DeviceGraph", cache_path: str | None = None, iterations: int = 10):
        """prh_autotune: time candidate launch plans on yg and keep the fastest as this context's tuning.
        Returns (Tuning, ms per iteration, whether the choice came from the cache)."""
        t = prh_tuning()
        ms = C.c_double()
        hit = C.c_int()
        path = cache_path.encode() if cache_path else None
        self.check(self._lib.prh_autotune(self._h, yg.handle, path, iterations, C.byref(t), C.byref(ms),
                                          C.byref(hit)))
        return _tuning_of(t), ms.value, bool(hit.value)

    def upload(self, y: HostGraph, with_out: bool = True) -> "DeviceGraph":
        return DeviceGraph(self, y, with_out)

    # --- kernel timing (HIP events around each rank-update launch) -----------------------------------------
    def kernel_timing(self, enable: bool):
        self.check(self._lib.prh_kernel_timing_enable(self._h, 1 if enable else 0))

    def kernel_timing_read(self) -> tuple[int, float]:
        n, ms = C.c_uint64(), C.c_double()
        self.check(self._lib.prh_kernel_timing_read(self._h, C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    # --- partition / marking (exposed for bit-exact checks) ------------------------------------------------
    def partition(self, offsets: np.ndarray, switch_degree: int = 64) -> tuple[np.ndarray, int]:
        """pagerankPartitionVerticesCuW (pagerankCuda.hxx:625-638): (order, partition point)."""
        off = np.ascontiguousarray(offsets, dtype=np.uint32)
        n = len(off) - 1
        order = np.zeros(max(n, 0), dtype=np.uint32)
        np_ = C.c_uint32()
        self.check(self._lib.prh_partition(self._h, n, off.ctypes.data_as(u32p), switch_degree,
                                           order.ctypes.data_as(u32p), C.byref(np_)))
        return order, int(np_.value)

    def mark_affected(self, yg: "DeviceGraph", method: int, batch: Batch) -> np.ndarray:
        out = np.zeros(yg.N, dtype=np.uint8)
        self.check(self._lib.prh_mark_affected(self._h, yg.handle, method, _p(batch.del_u), _p(batch.del_v),
                                               batch.deletions, _p(batch.ins_u), batch.insertions,
                                               out.ctypes.data_as(u8p)))
        return out


def _p(a: np.ndarray):
    return a.ctypes.data_as(u32p) if len(a) else None


class DeviceGraph:
    """A graph resident in device memory (prh_graph): y's in-CSR, out-degrees and, optionally, its out-CSR."""

    def __init__(self, ctx: Context, y: HostGraph, with_out: bool = True):
        self.ctx = ctx
        self.N, self.M = y.N, y.M
        h = C.c_void_p()
        oo = y.out_off.ctypes.data_as(u32p) if with_out else None
        oe = y.out_edg.ctypes.data_as(u32p) if with_out else None
        ctx.check(ctx._lib.prh_graph_create(ctx.handle, y.N, y.M, y.in_off.ctypes.data_as(u32p),
                                            y.in_edg.ctypes.data_as(u32p), y.outdeg.ctypes.data_as(u32p), oo, oe,
                                            C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def update(self, batch: Batch) -> float:
        """Apply the batch to this graph on the device (prh_graph_update): tidied as tidyBatchUpdateU tidies it, the
        propagation-blocked plan patched in place. Returns the update's milliseconds."""
        ms = C.c_float()
        self.ctx.check(self.ctx._lib.prh_graph_update(self.ctx.handle, self._h, _p(batch.del_u), _p(batch.del_v),
                                                      batch.deletions, _p(batch.ins_u), _p(batch.ins_v),
                                                      batch.insertions, C.byref(ms)))
        self.M = int(self.ctx._lib.prh_graph_size(self._h))
        return float(ms.value)

    def csr(self, which: str = "out") -> tuple[np.ndarray, np.ndarray]:
        """The graph's current CSR in vertex ids (prh_graph_csr): "in" (in-neighbours) or "out" (targets)."""
        off = np.zeros(self.N + 1, dtype=np.uint32)
        idx = np.zeros(max(self.M, 1), dtype=np.uint32)
        self.ctx.check(self.ctx._lib.prh_graph_csr(self.ctx.handle, self._h, 1 if which == "out" else 0,
                                                   off.ctypes.data_as(u32p), idx.ctypes.data_as(u32p)))
        return off, idx[:self.M]

    def close(self):
        if getattr(self, "_h", None) and self.ctx._h:
            self.ctx._lib.prh_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def solve(yg: DeviceGraph, method: int, o: PagerankOptions | None = None, batch: Batch | None = None,
          q: np.ndarray | None = None, partition: bool = True) -> PagerankResult:
    """One solver call on host buffers (prh_solve); the result mirrors PagerankResult."""
    o = o or PagerankOptions()
    b = batch or Batch.empty()
    ctx = yg.ctx
    ranks = np.zeros(yg.N, dtype=np.float64)
    qq = None
    if q is not None:
        qq = np.ascontiguousarray(q, dtype=np.float64)
        if len(qq) != yg.N:
            raise ValueError(f"initial ranks have {len(qq)} entries, graph has {yg.N} vertices")
    st = prh_stats()
    ctx.check(ctx._lib.prh_solve(ctx.handle, yg.handle, method, 1 if partition else 0, C.byref(o.native()),
                                 _p(b.del_u), _p(b.del_v), b.deletions, _p(b.ins_u), b.insertions,
                                 qq.ctypes.data_as(f64p) if qq is not None else None,
                                 ranks.ctypes.data_as(f64p), C.byref(st)))
    return PagerankResult(ranks, st.iterations, st.time, st.initialization_time, st.marking_time,
                          st.computation_time, st.preparation_time)


def solve_device(yg: DeviceGraph, method: int, o: PagerankOptions, q_ptr: int | None, ranks_ptr: int,
                 batch_ptrs: tuple[int, int, int, int, int] | None = None, partition: bool = True) -> prh_stats:
    """prh_solve_device on device pointers (e.g. torch tensors' data_ptr()); returns the native stats."""
    ctx = yg.ctx
    du, dv, nd, iu, ni = batch_ptrs or (None, None, 0, None, 0)
    st = prh_stats()
    ctx.check(ctx._lib.prh_solve_device(ctx.handle, yg.handle, method, 1 if partition else 0, C.byref(o.native()),
                                        du, dv, nd, iu, ni, q_ptr, ranks_ptr, C.byref(st)))
    return st


@dataclass
class BatchedResult:
    """Batched personalized PageRank: ranks[v, j] for column j, and each column's iteration count."""
    ranks: np.ndarray
    iterations: np.ndarray
    time: float = 0.0
    initializationTime: float = 0.0
    computationTime: float = 0.0


def ppr_batched(yg: DeviceGraph, seed_sets, o: PagerankOptions | None = None) -> BatchedResult:
    """Personalized PageRank for up to 64 teleport columns at once (prh_ppr_batched): column j teleports uniformly
    to seed_sets[j]."""
    o = o or PagerankOptions()
    k = len(seed_sets)
    ctx = yg.ctx
    sizes = [len(s) for s in seed_sets]
    offsets = np.zeros(k + 1, dtype=np.uint32)
    np.cumsum(sizes, out=offsets[1:])
    seeds = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.uint32) for s in seed_sets])
                                 if k else np.zeros(0, np.uint32))
    ranks = np.zeros((yg.N, max(k, 1)), dtype=np.float64)
    its = (C.c_int * max(k, 1))()
    st = prh_stats()
    ctx.check(ctx._lib.prh_ppr_batched(ctx.handle, yg.handle, k, offsets.ctypes.data_as(u32p),
                                       seeds.ctypes.data_as(u32p), C.byref(o.native()), ranks.ctypes.data_as(f64p),
                                       its, C.byref(st)))
    return BatchedResult(ranks[:, :k], np.array(its[:k]), st.time, st.initialization_time, st.computation_time)


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def weighted_pagerank(yg: DeviceGraph, weights: np.ndarray, teleport: np.ndarray,
                      o: PagerankOptions | None = None) -> PagerankResult:
    """Weighted PageRank (prh_weighted_pagerank) on host arrays: weights are FP32 in out-CSR edge order."""
    o = o or PagerankOptions()
    ctx = yg.ctx
    w = np.ascontiguousarray(weights, dtype=np.float32)
    s = np.ascontiguousarray(teleport, dtype=np.float64)
    if len(w) != yg.M or len(s) != yg.N:
        raise ValueError("weights need one value per edge and teleport one per vertex")
    r = np.zeros(yg.N)
    st = prh_stats()
    ctx.check(ctx._lib.prh_weighted_pagerank(ctx.handle, yg.handle, _ptr(w), _ptr(s), C.byref(o.native()), _ptr(r),
                                             C.byref(st), 0))
    return PagerankResult(r, st.iterations, st.time, st.initialization_time, st.marking_time, st.computation_time,
                          st.preparation_time)


def weighted_adjoint(yg: DeviceGraph, weights: np.ndarray, ranks: np.ndarray, grad_ranks: np.ndarray,
                     o: PagerankOptions | None = None):
    """Adjoint solve and gradient assembly (prh_weighted_adjoint) on host arrays.
    Returns (lambda, dL/dweights in out-CSR order, dL/dteleport, iterations)."""
    o = o or PagerankOptions()
    ctx = yg.ctx
    w = np.ascontiguousarray(weights, dtype=np.float32)
    r = np.ascontiguousarray(ranks, dtype=np.float64)
    g = np.ascontiguousarray(grad_ranks, dtype=np.float64)
    lam, gw, gs = np.zeros(yg.N), np.zeros(yg.M), np.zeros(yg.N)
    st = prh_stats()
    ctx.check(ctx._lib.prh_weighted_adjoint(ctx.handle, yg.handle, _ptr(w), _ptr(r), _ptr(g), C.byref(o.native()),
                                            _ptr(lam), _ptr(gw), _ptr(gs), C.byref(st), 0))
    return lam, gw, gs, st.iterations


class Stream:
    """Streaming dynamic session (prh_stream_*): a device copy of yg that batches update on the device, with each
    batch's solve (method from the previous ranks) replayed from captured HIP graphs."""

    def __init__(self, yg: DeviceGraph, method: int, o: PagerankOptions | None = None, max_batch: int = 1 << 16,
                 max_edges: int = 0, initial_ranks: np.ndarray | None = None):
        self.ctx = yg.ctx
        self.graph = yg  # keeps the base graph (its layout and permutation) alive
        self.N = yg.N
        o = o or PagerankOptions()
        q = None
        if initial_ranks is not None:
            q = np.ascontiguousarray(initial_ranks, dtype=np.float64)
            if len(q) != yg.N:
                raise ValueError("initial ranks need one value per vertex")
        h = C.c_void_p()
        self.ctx.check(self.ctx._lib.prh_stream_create(self.ctx.handle, yg.handle, method, C.byref(o.native()),
                                                       max_batch, max_edges,
                                                       q.ctypes.data_as(f64p) if q is not None else None,
                                                       C.byref(h)))
        self._h = h

    def update(self, batch: Batch, want_ranks: bool = True) -> PagerankResult:
        ranks = np.zeros(self.N) if want_ranks else np.zeros(0)
        st = prh_stats()
        self.ctx.check(self.ctx._lib.prh_stream_update(self._h, _p(batch.del_u), _p(batch.del_v), batch.deletions,
                                                       _p(batch.ins_u), _p(batch.ins_v), batch.insertions,
                                                       ranks.ctypes.data_as(f64p) if want_ranks else None,
                                                       C.byref(st)))
        return PagerankResult(ranks, st.iterations, st.time, st.initialization_time, st.marking_time,
                              st.computation_time)

    def ranks(self) -> np.ndarray:
        r = np.zeros(self.N)
        self.ctx.check(self.ctx._lib.prh_stream_ranks(self._h, r.ctypes.data_as(f64p)))
        return r

    @property
    def size(self) -> int:
        return int(self.ctx._lib.prh_stream_size(self._h))

    def csr(self, which: str = "out") -> tuple[np.ndarray, np.ndarray]:
        """The session's current CSR in vertex ids (prh_stream_csr): "in" or "out"."""
        m = self.size
        off = np.zeros(self.N + 1, dtype=np.uint32)
        idx = np.zeros(max(m, 1), dtype=np.uint32)
        self.ctx.check(self.ctx._lib.prh_stream_csr(self._h, 1 if which == "out" else 0, off.ctypes.data_as(u32p),
                                                    idx.ctypes.data_as(u32p)))
        return off, idx[:m]

    def close(self):
        if getattr(self, "_h", None):
            self.ctx._lib.prh_stream_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Reference-named wrappers (pagerankCuda.hxx:864-988). x is accepted for signature parity and, like the
# reference, only its emptiness matters.
def pagerankStaticHip(yg: DeviceGraph, o: PagerankOptions | None = None, partition: bool = True) -> PagerankResult:
    return solve(yg, STATIC, o, None, None, partition)


def pagerankNaiveDynamicHip(yg: DeviceGraph, q, o: PagerankOptions | None = None,
                            partition: bool = True) -> PagerankResult:
    return solve(yg, NAIVE_DYNAMIC, o, None, q, partition)


def pagerankDynamicTraversalHip(yg: DeviceGraph, batch: Batch, q, o: PagerankOptions | None = None,
                                partition: bool = True) -> PagerankResult:
    return solve(yg, DYNAMIC_TRAVERSAL, o, batch, q, partition)


def pagerankDynamicFrontierHip(yg: DeviceGraph, batch: Batch, q, o: PagerankOptions | None = None,
                               partition: bool = True) -> PagerankResult:
    return solve(yg, DYNAMIC_FRONTIER, o, batch, q, partition)


def pagerankPruneDynamicFrontierHip(yg: DeviceGraph, batch: Batch, q, o: PagerankOptions | None = None,
                                    partition: bool = True) -> PagerankResult:
    return solve(yg, PRUNE_DYNAMIC_FRONTIER, o, batch, q, partition)
