"""The propagation-blocked plan of the plain update (tuning.blocking) against the oracle, through the C ABI.

It changes only the summation order (LDS additions in arrival order), so the bars are the row plan's: ranks within
L-inf 1e-9 of the oracle's restatement of the reference GPU path, iterations equal or +-1. The graphs cover one
bin smaller than B, several bins with a partial last one, bins split over several items (items hold at least 16K
slots: RMAT-14's heaviest 1024 rows hold 148K in-edges, finished a thread per vertex; RMAT-15's first 16384 rows
hold more than 16 items' worth, finished 16 waves per 64 vertices), rows without in-edges, one destination for
all slots of a group (the star), and groups of 64 slots spanning several short runs (a sparse uniform graph: about
ten edges per nonempty off-diagonal (block, bin) tile at B = 1024, so most groups hold more than two runs and their
later lanes take their X offset from the run table rather than from the group table).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import star_graph

pytestmark = pytest.mark.gpu

RANK_TOL = 1e-9


def _with_blocking(ctx, blocking, fn):
    from pagerank_amd import Tuning
    old = ctx.get_tuning()
    t = Tuning(**{**old.__dict__, "blocking": blocking})
    try:
        ctx.set_tuning(t)
        return fn()
    finally:
        ctx.set_tuning(old)


def _graph(name):
    from pagerank_amd import HostGraph
    if name == "rmat14":
        return HostGraph.rmat(14, 16, seed=11)
    if name == "rmat15":
        return HostGraph.rmat(15, 16, seed=3)
    if name == "rmat12x24":
        return HostGraph.rmat(12, 24, seed=4)
    if name == "star5000":
        return star_graph(5000)
    if name == "grid70x50":
        return HostGraph.grid(70, 50, 0.62, seed=5)
    if name == "uniform200k":
        n, m = 200_000, 400_000
        rng = np.random.default_rng(8)
        return HostGraph.from_edges(n, rng.integers(0, n, m).astype(np.uint32), rng.integers(0, n, m).astype(np.uint32),
                                    self_loops=True)
    n = 3000  # rows without in-edges, no self-loops
    u = np.arange(n, dtype=np.uint32)
    src = np.concatenate([u[1:], [0], u[:100]])
    dst = np.concatenate([np.zeros(n - 1, np.uint32), [1], (u[:100] * 7 + 3) % 50])
    return HostGraph.from_edges(n, src, dst, self_loops=False)


CASES = [("rmat14", 1024), ("rmat14", 4096), ("rmat12x24", 2048), ("star5000", 8192), ("grid70x50", 1024),
         ("no_in_edges", 1024), ("rmat14", 16384), ("rmat15", 16384), ("uniform200k", 1024)]


@pytest.mark.parametrize("name,blocking", CASES, ids=[f"{a}_B{b}" for a, b in CASES])
def test_binned_static_matches_oracle(ctx, oracle_mod, name, blocking):
    from pagerank_amd import pagerankStaticHip
    y = _graph(name)
    print(f"[binned] graph={name} B={blocking} N={y.N} M={y.M} max_in={int(np.diff(y.in_off).max())}", flush=True)
    res = _with_blocking(ctx, blocking, lambda: pagerankStaticHip(ctx.upload(y)))
    ref, it, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y)
    assert abs(res.iterations - it) <= 1, (res.iterations, it)
    assert np.abs(res.ranks - ref).max() <= RANK_TOL


def test_binned_naive_dynamic(ctx, oracle_mod):
    """Naive-dynamic runs the same plain update from the previous ranks."""
    from pagerank_amd import Batch, HostGraph, apply_batch, solve
    x = HostGraph.rmat(13, 16, seed=9)
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    b = Batch.random(x, 300, 0.8, seed=2)
    y = apply_batch(x, b)
    print(f"[binned-nd] N={y.N} M={y.M}", flush=True)
    res = _with_blocking(ctx, 2048, lambda: solve(ctx.upload(y), 1, batch=b, q=q))
    ref, it, _ = oracle_mod.pagerank_cuda(1, y, b, q)
    assert abs(res.iterations - it) <= 1, (res.iterations, it)
    assert np.abs(res.ranks - ref).max() <= RANK_TOL


DYN = [(2, 1024, 1e-3), (3, 2048, 1e-3), (4, 1024, 1e-3), (3, 4096, 1e-2), (4, 2048, 1e-4), (4, 8192, 1e-4),
       (2, 16384, 1e-3)]


@pytest.mark.parametrize("method,blocking,frac", DYN, ids=[f"m{m}_B{b}_{f:g}" for m, b, f in DYN])
def test_binned_dynamic_methods(ctx, oracle_mod, method, blocking, frac):
    """DT, DF and DF-P through the binned plan: every vertex gets its full in-sum, unaffected ones are carried over,
    prune and frontier flags as the row kernel sets them, extension direction chosen from the costs the binned
    kernels record. Ranks within 1e-9 of the oracle's restatement of the same method, iterations +-1."""
    from pagerank_amd import Batch, HostGraph, apply_batch, solve
    x = HostGraph.rmat(13, 16, seed=21)
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    b = Batch.random(x, max(1, int(frac * x.M)), 0.8, seed=5)
    y = apply_batch(x, b)
    print(f"[binned-dyn] method={method} B={blocking} frac={frac} N={y.N} M={y.M}", flush=True)
    res = _with_blocking(ctx, blocking, lambda: solve(ctx.upload(y), method, batch=b, q=q))
    ref, it, _ = oracle_mod.pagerank_cuda(method, y, b, q)
    assert abs(res.iterations - it) <= 1, (res.iterations, it)
    assert np.abs(res.ranks - ref).max() <= RANK_TOL


def test_binned_repeats_to_rounding(ctx):
    """Two solves with the plan agree to rounding (LDS additions may arrive in another order)."""
    from pagerank_amd import HostGraph, pagerankStaticHip
    y = HostGraph.rmat(14, 16, seed=11)
    yg = ctx.upload(y)
    a = _with_blocking(ctx, 1024, lambda: pagerankStaticHip(yg))
    b = _with_blocking(ctx, 1024, lambda: pagerankStaticHip(yg))
    assert abs(a.iterations - b.iterations) <= 1
    assert np.abs(a.ranks - b.ranks).max() <= 1e-12
