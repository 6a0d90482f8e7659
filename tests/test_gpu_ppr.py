"""Batched personalized PageRank on the GPU (prh_ppr_batched) against the oracle's single-column solve.

Bars: every column within L-inf 1e-9 of orc_ppr (the FP64 single-column solve with the same stopping rule) and with
the same iteration count +-1; a column whose teleport is uniform over all vertices equals the static solve within
1e-9; every column sums to 1 (no dead ends); results do not depend on how many iterations one HIP-graph replay
holds (the captured run is replayed many times and checked every time against the eager CPU result).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import star_graph

pytestmark = pytest.mark.gpu


def _seed_sets(n, k, max_size, rng):
    return [rng.choice(n, size=int(rng.integers(1, max_size + 1)), replace=False).astype(np.uint32) for _ in range(k)]


def _teleport(n, seeds):
    s = np.zeros(n)
    np.add.at(s, seeds, 1.0 / len(seeds))
    return s


@pytest.mark.parametrize("scale,k", [(11, 64), (13, 64), (12, 5)])
def test_batched_matches_single_column_oracle(ctx, oracle_mod, scale, k):
    from pagerank_amd import HostGraph, ppr_batched
    y = HostGraph.rmat(scale, 16, seed=scale + k)
    rng = np.random.default_rng(scale * 100 + k)
    seeds = _seed_sets(y.N, k, min(1000, y.N // 4), rng)
    print(f"[ppr] scale={scale} k={k} N={y.N} M={y.M}", flush=True)
    res = ppr_batched(ctx.upload(y, with_out=False), seeds)
    assert res.ranks.shape == (y.N, k)
    for j in range(k):
        s = _teleport(y.N, seeds[j])
        ref, it = oracle_mod.ppr(y, s)
        assert abs(int(res.iterations[j]) - it) <= 1, (j, res.iterations[j], it)
        assert np.abs(res.ranks[:, j] - ref).max() <= 1e-9, j
        assert abs(res.ranks[:, j].sum() - 1.0) < 1e-9


def test_uniform_column_equals_static(ctx, oracle_mod):
    from pagerank_amd import HostGraph, pagerankStaticHip, ppr_batched
    y = HostGraph.rmat(12, 16, seed=3)
    yg = ctx.upload(y, with_out=False)
    seeds = [np.arange(y.N, dtype=np.uint32), np.array([7], dtype=np.uint32)]
    res = ppr_batched(yg, seeds)
    st = pagerankStaticHip(yg)
    assert abs(int(res.iterations[0]) - st.iterations) <= 1
    assert np.abs(res.ranks[:, 0] - st.ranks).max() <= 1e-9


def test_heavy_vertex_chunks(ctx, oracle_mod):
    """A hub with in-degree above the chunk size is summed in chunks by many waves and combined."""
    from pagerank_amd import ppr_batched
    y = star_graph(6000)
    seeds = [np.array([0], np.uint32), np.array([1, 2, 3], np.uint32), np.arange(1, 6000, 7, dtype=np.uint32)]
    res = ppr_batched(ctx.upload(y, with_out=False), seeds)
    for j, sd in enumerate(seeds):
        ref, it = oracle_mod.ppr(y, _teleport(y.N, sd))
        assert abs(int(res.iterations[j]) - it) <= 1
        assert np.abs(res.ranks[:, j] - ref).max() <= 1e-9


@pytest.mark.parametrize("interval", [1, 3, 16])
def test_graph_replay_interval_does_not_change_results(ctx, interval):
    from pagerank_amd import HostGraph, Tuning, ppr_batched
    y = HostGraph.rmat(11, 16, seed=9)
    yg = ctx.upload(y, with_out=False)
    seeds = _seed_sets(y.N, 64, 100, np.random.default_rng(1))
    old = ctx.get_tuning()
    try:
        ctx.set_tuning(Tuning(old.switch_degree, old.huge_degree, old.low_width, 8))
        base = ppr_batched(yg, seeds)
        ctx.set_tuning(Tuning(old.switch_degree, old.huge_degree, old.low_width, interval))
        for _ in range(2):  # replays of a freshly captured graph, twice, with eager work in between
            again = ppr_batched(yg, seeds)
            assert np.array_equal(again.ranks, base.ranks)
            assert np.array_equal(again.iterations, base.iterations)
    finally:
        ctx.set_tuning(old)


def test_bad_seeds_rejected(ctx):
    from pagerank_amd import HostGraph, PagerankError, ppr_batched
    yg = ctx.upload(HostGraph.rmat(8, 4, seed=1), with_out=False)
    with pytest.raises(PagerankError):
        ppr_batched(yg, [np.array([10**6], np.uint32)])
    with pytest.raises(PagerankError):
        ppr_batched(yg, [np.zeros(0, np.uint32)])
    with pytest.raises(PagerankError):
        ppr_batched(yg, [np.array([1], np.uint32)] * 65)


def test_uncaptured_mode_matches_captured(monkeypatch):
    """PRH_GRAPH_CAPTURE=0 at context creation: the batched solve and the streaming session issue their launches
    one by one instead of replaying captured graphs; the results are the same bits (no order-dependent sums here:
    the row plan and the batched solver sum in a fixed order)."""
    from pagerank_amd import Batch, Context, HostGraph, Stream, engine
    y = HostGraph.rmat(12, 16, seed=8)
    rng = np.random.default_rng(2)
    sets = [rng.choice(y.N, size=int(rng.integers(1, 50)), replace=False).astype(np.uint32) for _ in range(5)]
    b = Batch.random(y, 60, 0.8, seed=3)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("PRH_GRAPH_CAPTURE", mode)
        with Context(0) as c:
            yg = c.upload(y)
            res = engine.ppr_batched(yg, sets)
            st = Stream(yg, 4, max_batch=128)
            got = st.update(b)
            st.close()
            yg.close()
        out[mode] = (res.ranks, res.iterations, got.ranks, got.iterations)
    assert np.array_equal(out["0"][0], out["1"][0]) and np.array_equal(out["0"][1], out["1"][1])
    assert np.array_equal(out["0"][2], out["1"][2]) and out["0"][3] == out["1"][3]
