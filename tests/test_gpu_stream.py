"""The streaming session (prh_stream_*): batches applied on the device, each solved from the previous ranks.

Bars: after every batch the device graph has the edge count of the host-applied graph; the ranks equal those of
the same method called on the host-applied graph with the previous batch's ranks (a separate upload, so a
different summation order) to L-inf 1e-9; batch tidying follows tidyBatchUpdateU; bad batches fail loudly.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
RANK_TOL = 1e-9


def _sequence(x, n, size, seed):
    from pagerank_amd import Batch, apply_batch
    out = []
    for i in range(n):
        b = Batch.random(x, size, 0.8, seed=seed + i)
        y = apply_batch(x, b)
        out.append((b, y))
        x = y
    return out


@pytest.mark.parametrize("blocking", [0, 1024], ids=["rows", "blocked"])
@pytest.mark.parametrize("method", [4, 3, 2, 1, 0], ids=["DFP", "DF", "DT", "ND", "static"])
def test_stream_matches_sequential_calls(ctx, oracle_mod, method, blocking):
    """16 consecutive batches through the stream equal 16 sequential calls of the same method on host-built graphs;
    after every batch the session's in-CSR and out-CSR equal the host-applied graph's bit for bit. blocked: the
    session's update runs through its own propagation-blocked plan (B = 1024), patched per batch."""
    from pagerank_amd import HostGraph, Stream, Tuning, solve
    x = HostGraph.rmat(12, 16, seed=31)
    q0, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    seq = _sequence(x, 16, max(1, int(1e-3 * x.M)), 100)
    old = ctx.get_tuning()
    ctx.set_tuning(Tuning(**{**old.__dict__, "blocking": blocking}))
    try:
        xg = ctx.upload(x)
        st = Stream(xg, method, max_batch=4096, initial_ranks=q0)
        q = q0
        for i, (b, y) in enumerate(seq):
            print(f"[stream] method={method} blocking={blocking} batch={i} -{b.deletions}/+{b.insertions}", flush=True)
            got = st.update(b)
            assert st.size == y.M
            for which, off, idx in (("out", y.out_off, y.out_edg), ("in", y.in_off, y.in_edg)):
                go, gi = st.csr(which)
                assert np.array_equal(go, off) and np.array_equal(gi, idx), (i, which)
            ref = solve(ctx.upload(y), method, batch=b, q=q if method else None)
            assert abs(got.iterations - ref.iterations) <= 1, (i, got.iterations, ref.iterations)
            assert np.abs(got.ranks - ref.ranks).max() <= RANK_TOL, i
            q = ref.ranks
        if method in (1, 2, 3):  # exact methods: the last ranks are the fixed point of the last graph
            exact, _ = oracle_mod.power_iteration(seq[-1][1])
            assert np.abs(st.ranks() - exact).max() <= RANK_TOL
        st.close()
    finally:
        ctx.set_tuning(old)


def test_stream_traversal_sets_match_oracle(ctx, oracle_mod):
    """DT through the stream: after each batch the ranks equal the oracle's DT solve on the host-applied graph, whose
    marking is the traversal closure (bit-exact with the reference's BFS closure, tests/test_oracle.py)."""
    from pagerank_amd import HostGraph, Stream
    x = HostGraph.rmat(11, 8, seed=17)
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    st = Stream(ctx.upload(x), 2, max_batch=512, initial_ranks=q)
    for i, (b, y) in enumerate(_sequence(x, 6, 40, 900)):
        got = st.update(b)
        ref, it, _ = oracle_mod.pagerank_cuda(2, y, b, q)
        assert abs(got.iterations - it) <= 1 and np.abs(got.ranks - ref).max() <= RANK_TOL, i
        q = got.ranks
    st.close()


def test_stream_tidies_batches(ctx, oracle_mod):
    """Duplicates, deletions of absent edges, insertions of present edges and an edge both deleted and inserted are
    tidied as tidyBatchUpdateU does; the resulting graph and ranks equal those of the host-tidied batch."""
    from pagerank_amd import Batch, HostGraph, Stream, apply_batch, solve
    x = HostGraph.rmat(10, 8, seed=3)
    q0, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    rng = np.random.default_rng(4)
    src = x.out_sources()
    nsl = np.flatnonzero(src != x.out_edg)
    e = rng.choice(nsl, 40, replace=False)
    dels = np.stack([src[e], x.out_edg[e]], 1)
    dels = np.concatenate([dels, dels[:5], [[1, 2], [3, 4]]])  # duplicates and probably-absent edges
    ins = rng.integers(0, x.N, size=(60, 2))
    ins = np.concatenate([ins, ins[:7], dels[:3], np.stack([src[:4], x.out_edg[:4]], 1)])  # dup, del+ins, present
    raw = Batch(np.ascontiguousarray(dels[:, 0], np.uint32), np.ascontiguousarray(dels[:, 1], np.uint32),
                np.ascontiguousarray(ins[:, 0], np.uint32), np.ascontiguousarray(ins[:, 1], np.uint32))
    tidy = Batch.tidy(x, dels, ins)
    y = apply_batch(x, tidy)
    st = Stream(ctx.upload(x), 4, max_batch=256, initial_ranks=q0)
    got = st.update(raw)
    assert st.size == y.M
    ref = solve(ctx.upload(y), 4, batch=tidy, q=q0)
    assert np.abs(got.ranks - ref.ranks).max() <= RANK_TOL
    st.close()


def test_stream_rejects_bad_batches(ctx):
    from pagerank_amd import Batch, HostGraph, PagerankError, Stream
    x = HostGraph.rmat(9, 8, seed=5)
    z = np.zeros(0, np.uint32)
    st = Stream(ctx.upload(x), 4, max_batch=8)
    with pytest.raises(PagerankError, match="capacity"):
        st.update(Batch(z, z, np.arange(9, dtype=np.uint32), np.arange(9, dtype=np.uint32)))
    st.update(Batch.empty())  # still usable: nothing was applied
    with pytest.raises(PagerankError, match="out of range"):
        st.update(Batch(z, z, np.array([0], np.uint32), np.array([x.N + 5], np.uint32)))
    with pytest.raises(PagerankError, match="unusable"):
        st.update(Batch.empty())
    st.close()
    # a batch deleting a vertex's only out-edge (its self-loop) would leave a dead end
    y = HostGraph.from_edges(4, np.array([0, 1, 2], np.uint32), np.array([1, 2, 0], np.uint32), self_loops=True)
    st = Stream(ctx.upload(y), 1, max_batch=8)
    with pytest.raises(PagerankError, match="dead end"):
        st.update(Batch(np.array([3], np.uint32), np.array([3], np.uint32), z, z))
    st.close()


def test_stream_on_fresh_context(oracle_mod):
    """A context whose first dynamic work is a stream (no earlier dynamic solve has sized its workspace)."""
    from pagerank_amd import Batch, Context, HostGraph, Stream, apply_batch, solve
    x = HostGraph.rmat(11, 16, seed=9)
    b = Batch.random(x, 50, 0.8, seed=1)
    y = apply_batch(x, b)
    with Context(0) as fresh:
        xg = fresh.upload(x)
        q = solve(xg, 0).ranks  # a static solve only
        st = Stream(xg, 4, max_batch=64, initial_ranks=q)
        got = st.update(b)
        st.close()
        ref, it, _ = oracle_mod.pagerank_cuda(oracle_mod.PRUNE_DYNAMIC_FRONTIER, y, b, q)
        assert abs(got.iterations - it) <= 1 and np.abs(got.ranks - ref).max() <= RANK_TOL
