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


def test_dead_end_rejected(ctx):
    from pagerank_amd import HostGraph, PagerankError
    y = HostGraph.from_edges(4, [0, 1, 2], [1, 2, 3], self_loops=False)  # vertex 3 has out-degree 0
    with pytest.raises(PagerankError) as e:
        ctx.upload(y)
    assert e.value.status == 2 and "dead-end" in str(e.value)


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
