"""Failure handling through the C ABI, forced from the host: every case returns a status (never exits) and the
GPU only runs kernels that exit normally."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_empty_graph_returns_empty_result(ctx):
    from pagerank_amd import HostGraph, pagerankStaticHip
    y = HostGraph(arrays=dict(out_off=np.zeros(1), out_edg=np.zeros(0), in_off=np.zeros(1), in_edg=np.zeros(0),
                              in_eid=np.zeros(0), outdeg=np.zeros(0)))
    res = pagerankStaticHip(ctx.upload(y))
    assert res.iterations == 0 and len(res.ranks) == 0


def _dead_end_graph():
    """RMAT-11 edges without self-loops, where a quarter of the vertices keep no out-edge (dead ends)."""
    from pagerank_amd import HostGraph
    x = HostGraph.rmat(11, 8, seed=31, self_loops=False)
    src = np.repeat(np.arange(x.N, dtype=np.uint32), np.diff(x.out_off.astype(np.int64)))
    keep = (src % 4) != 3
    return HostGraph.from_edges(x.N, src[keep], x.out_edg[keep], self_loops=False)


def test_dead_ends_accepted_except_dfp(ctx, oracle_mod):
    """Dead ends (out-degree 0) are accepted as the reference accepts them: a dead end is never an in-neighbour, so
    nothing divides by its degree. Static, ND, DT and DF match the oracle; DF-P, whose closed form divides by the
    out-degree, is rejected with a status."""
    from pagerank_amd import (DYNAMIC_FRONTIER, DYNAMIC_TRAVERSAL, NAIVE_DYNAMIC, PRUNE_DYNAMIC_FRONTIER, STATIC,
                              Batch, PagerankError, apply_batch, solve)
    x = _dead_end_graph()
    assert int((x.outdeg == 0).sum()) > 100
    b = Batch.random(x, 40, 0.8, seed=5)
    y = apply_batch(x, b)
    yg = ctx.upload(y)
    q, _, _ = oracle_mod.pagerank_cuda(STATIC, x)
    for m in (STATIC, NAIVE_DYNAMIC, DYNAMIC_TRAVERSAL, DYNAMIC_FRONTIER):
        res = solve(yg, m, batch=b, q=None if m == STATIC else q)
        ref, it, _ = oracle_mod.pagerank_cuda(m, y, b, None if m == STATIC else q)
        assert abs(res.iterations - it) <= 1, m
        assert np.abs(res.ranks - ref).max() <= 1e-9, m
        assert res.ranks.sum() < 1.0  # rank mass leaks through the dead ends
    with pytest.raises(PagerankError) as e:
        solve(yg, PRUNE_DYNAMIC_FRONTIER, batch=b, q=q)
    assert e.value.status == 2 and "dead ends" in str(e.value)


def test_out_of_range_id_rejected(ctx):
    from pagerank_amd import HostGraph, PagerankError
    y = HostGraph.rmat(6, 4, seed=1)
    bad = HostGraph(arrays=dict(out_off=y.out_off, out_edg=y.out_edg, in_off=y.in_off,
                                in_edg=np.where(np.arange(y.M) == 5, 10**6, y.in_edg), in_eid=y.in_eid,
                                outdeg=y.outdeg))
    with pytest.raises(PagerankError) as e:
        ctx.upload(bad)
    assert e.value.status == 2


def test_out_of_range_batch_rejected(ctx):
    from pagerank_amd import Batch, HostGraph, PagerankError, solve
    y = HostGraph.rmat(6, 4, seed=1)
    b = Batch(np.array([0], np.uint32), np.array([10**6], np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    with pytest.raises(PagerankError) as e:
        solve(ctx.upload(y), 3, batch=b, q=np.full(y.N, 1.0 / y.N))
    assert e.value.status == 1


def test_nan_initial_ranks_rejected(ctx):
    from pagerank_amd import HostGraph, PagerankError, pagerankNaiveDynamicHip
    y = HostGraph.rmat(6, 4, seed=1)
    q = np.full(y.N, 1.0 / y.N)
    q[3] = np.nan
    with pytest.raises(PagerankError) as e:
        pagerankNaiveDynamicHip(ctx.upload(y), q)
    assert e.value.status == 1


def test_zero_tolerance_runs_max_iterations(ctx):
    from pagerank_amd import HostGraph, PagerankOptions, pagerankStaticHip
    y = HostGraph.rmat(8, 8, seed=1)
    res = pagerankStaticHip(ctx.upload(y), PagerankOptions(tolerance=0.0, maxIterations=37))
    assert res.iterations == 37


def test_dynamic_without_out_csr_unsupported(ctx):
    from pagerank_amd import Batch, HostGraph, PagerankError, solve
    y = HostGraph.rmat(6, 4, seed=1)
    yg = ctx.upload(y, with_out=False)
    with pytest.raises(PagerankError) as e:
        solve(yg, 3, batch=Batch.empty(), q=np.full(y.N, 1.0 / y.N))
    assert e.value.status == 5


def test_memory_limit_gives_out_of_memory():
    from pagerank_amd import Context, HostGraph, PagerankError
    y = HostGraph.rmat(10, 8, seed=1)
    with Context(0) as c:
        c.set_memory_limit(1024)
        with pytest.raises(PagerankError) as e:
            c.upload(y)
        assert e.value.status == 3
        assert c.memory_used() == 0


def test_bad_options_rejected(ctx):
    from pagerank_amd import HostGraph, PagerankError, PagerankOptions, pagerankStaticHip
    yg = ctx.upload(HostGraph.rmat(6, 4, seed=1))
    for o in (PagerankOptions(repeat=0), PagerankOptions(damping=1.0), PagerankOptions(maxIterations=-1),
              PagerankOptions(tolerance=float("nan"))):
        with pytest.raises(PagerankError) as e:
            pagerankStaticHip(yg, o)
        assert e.value.status == 1
