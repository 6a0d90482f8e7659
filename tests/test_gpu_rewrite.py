"""The device CSR rewrite of an update (rewrite_kernel: tiles of 16384 old edges) in every form it takes.

A tile over a few units (runs of unchanged rows, each closed by a changed row) moves them unit by unit: the
unchanged rows as one shifted run, the changed row against its deletions and insertions staged in LDS (up to 256 of
each) or, past that, searched in the batch's keys. A tile over many units moves edge by edge. The cases:
- an RMAT graph of 2^16 vertices (64 tiles) with batches of 20 edges: tiles of one or two units, hub rows across
  tile boundaries;
- a hub whose in-row and out-row each span three tiles, with 50 (staged) or 300 (searched) deletions and as many
  insertions in each;
- a ring of 30,000 vertices (three edges a row) with batches of 500 and 5,000 random edges: tiles over tens of
  units (their units and row ranges staged in LDS) and over more than 256 (searched in memory);
- an RMAT graph of 2^13 vertices with a batch of 0.2% of its edges (tens of units per tile), covered by
  test_gpu_update.py.

Bar: after every batch the device graph's in-CSR and out-CSR, read back in vertex ids, equal the host-applied graph
(apply_batch, pinned to the reference's applyBatchUpdateOmpU) bit for bit, through prh_graph_update and through the
streaming session.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hub_graph(n):
    """Vertex 0 receives an edge from every even vertex and sends one to every odd vertex; a ring joins the rest."""
    from pagerank_amd import HostGraph
    v = np.arange(1, n, dtype=np.uint32)
    even, odd = v[v % 2 == 0], v[v % 2 == 1]
    src = np.concatenate([even, np.zeros(len(odd), np.uint32), v])
    dst = np.concatenate([np.zeros(len(even), np.uint32), odd, np.where(v + 1 < n, v + 1, 1).astype(np.uint32)])
    return HostGraph.from_edges(n, src, dst)


def _hub_batch(x, k, part):
    """k of the hub's in-edges and k of its out-edges deleted, k absent edges of each side inserted, one ring edge
    deleted. Batch `part` (0 or 1) draws from every other candidate, so two batches never touch the same edge."""
    from pagerank_amd import Batch
    rng = np.random.default_rng(5 + part)
    n = x.N
    even = np.arange(2, n, 2, dtype=np.uint32)[part::2]
    odd = np.arange(1, n, 2, dtype=np.uint32)[part::2]
    d_in, i_out = np.split(rng.choice(even, 2 * k, replace=False), 2)
    d_out, i_in = np.split(rng.choice(odd, 2 * k, replace=False), 2)
    w = 1001 + 2 * part
    dels = [(int(u), 0) for u in d_in] + [(0, int(u)) for u in d_out] + [(w, w + 1)]
    ins = [(int(u), 0) for u in i_in] + [(0, int(u)) for u in i_out]
    return Batch.tidy(x, dels, ins)


def _check(dev, y, tag):
    for which, off, idx in (("out", y.out_off, y.out_edg), ("in", y.in_off, y.in_edg)):
        go, gi = dev.csr(which)
        assert np.array_equal(go, off), (tag, which, "offsets")
        assert np.array_equal(gi, idx), (tag, which, "indices")


def _batches(kind):
    from pagerank_amd import Batch, HostGraph, apply_batch
    if kind == "rmat16":
        x = HostGraph.rmat(16, 16, seed=23)
        seq, y = [], x
        for i in range(3):
            b = Batch.random(y, 20, 0.8, seed=70 + i)
            y = apply_batch(y, b)
            seq.append((b, y))
        return x, seq
    if kind == "ring":
        n = 30_000
        v = np.arange(n, dtype=np.uint32)
        x = HostGraph.from_edges(n, np.concatenate([v, v]), np.concatenate([(v + 1) % n, (v + 7) % n]))
        seq, y = [], x
        for i, size in enumerate((500, 5000)):
            b = Batch.random(y, size, 0.8, seed=60 + i)
            y = apply_batch(y, b)
            seq.append((b, y))
        return x, seq
    k = 50 if kind == "hub-staged" else 300
    x = _hub_graph(100_000)  # the hub's rows: 50,000 edges each, over three or four tiles
    seq, y = [], x
    for i in range(2):
        b = _hub_batch(y, k, i)
        assert b.deletions == 2 * k + 1 and b.insertions == 2 * k
        y = apply_batch(y, b)
        seq.append((b, y))
    return x, seq


KINDS = ["rmat16", "hub-staged", "hub-searched", "ring"]


@pytest.mark.parametrize("kind", KINDS)
def test_graph_update_rewrite(kind):
    from pagerank_amd import Context
    x, seq = _batches(kind)
    with Context(0) as c:
        xg = c.upload(x)
        for i, (b, y) in enumerate(seq):
            print(f"[rewrite] update {kind} batch={i} -{b.deletions}/+{b.insertions} M={y.M}", flush=True)
            xg.update(b)
            assert xg.M == y.M
            _check(xg, y, (kind, i))
        xg.close()


@pytest.mark.parametrize("kind", KINDS)
def test_stream_rewrite(ctx, kind):
    from pagerank_amd import Stream
    x, seq = _batches(kind)
    xg = ctx.upload(x)
    st = Stream(xg, 1, max_batch=8192)
    for i, (b, y) in enumerate(seq):
        print(f"[rewrite] stream {kind} batch={i} -{b.deletions}/+{b.insertions} M={y.M}", flush=True)
        st.update(b, want_ranks=False)
        assert st.size == y.M
        _check(st, y, (kind, i))
    st.close()
    xg.close()


def test_update_counts_dead_ends():
    """After an update the graph's vertices without out-edges are counted again (DF-P refuses them, naming the
    count); a later batch that gives them out-edges back lifts the refusal."""
    from pagerank_amd import Batch, Context, HostGraph, PagerankError, apply_batch, solve
    n = 300_000
    v = np.arange(n, dtype=np.uint32)
    x = HostGraph.from_edges(n, np.concatenate([v, v]), np.concatenate([(v + 1) % n, (v + 2) % n]), self_loops=False)
    ends = [5, 150_001, n - 10]
    cut = Batch.tidy(x, [(u, (u + 1) % n) for u in ends] + [(u, (u + 2) % n) for u in ends], [])
    y = apply_batch(x, cut)
    back = Batch.tidy(y, [], [(u, (u + 3) % n) for u in ends])
    z = apply_batch(y, back)
    q = np.full(n, 1.0 / n)
    with Context(0) as c:
        xg = c.upload(x)
        xg.update(cut)
        _check(xg, y, "cut")
        with pytest.raises(PagerankError) as e:
            solve(xg, 4, batch=cut, q=q)
        assert f"{len(ends)} vertices have none" in str(e.value)
        assert solve(xg, 1, batch=cut, q=q).iterations > 0  # the other methods accept dead ends
        xg.update(back)
        _check(xg, z, "back")
        assert solve(xg, 4, batch=back, q=q).iterations > 0
        xg.close()
