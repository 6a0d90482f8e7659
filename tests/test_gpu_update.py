"""prh_graph_update: batches applied to an uploaded graph on the device, its propagation-blocked plan patched in
place (dead slots and an overlay of inserted edges) instead of rebuilt.

Bars: after every batch the graph's in-CSR and out-CSR, read back in vertex ids, equal the host-applied graph
(apply_batch, pinned to the reference's applyBatchUpdateOmpU by tests/test_oracle.py) bit for bit; the DT affected set
is bit-exact with the oracle's; every solver on the updated graph matches the oracle on the host-applied graph
(L-inf <= 1e-9, iterations within 1; DF-P against its own restatement). The series includes a batch that re-inserts
deleted edges (clearing dead flags) and deletes inserted ones (removing overlay entries).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _series(x, n, size, seed):
    """n batches in a row: random ones, then one that undoes part of the earlier batches."""
    from pagerank_amd import Batch, apply_batch
    out, y, ins, dels = [], x, [], []
    for i in range(n - 1):
        b = Batch.random(y, size, 0.8, seed=seed + i)
        ins += list(zip(b.ins_u.tolist(), b.ins_v.tolist()))
        dels += list(zip(b.del_u.tolist(), b.del_v.tolist()))
        y = apply_batch(y, b)
        out.append((b, y))
    undo = Batch.tidy(y, ins[::3], dels[::2])  # delete inserted edges, re-insert deleted ones
    assert undo.deletions > 0 and undo.insertions > 0
    out.append((undo, apply_batch(y, undo)))
    return out


@pytest.mark.parametrize("blocking", [0, 1024, 2048], ids=["rows", "blocked1024", "blocked2048"])
def test_update_series_matches_host_applied_graph(oracle_mod, blocking):
    from pagerank_amd import Context, HostGraph, Tuning, solve
    x = HostGraph.rmat(13, 16, seed=9)
    seq = _series(x, 6, max(1, int(2e-3 * x.M)), 40)
    with Context(0) as c:
        c.set_tuning(Tuning(**{**c.get_tuning().__dict__, "blocking": blocking}))
        xg = c.upload(x)
        q, _, _ = oracle_mod.pagerank_cuda(0, x)
        r0 = solve(xg, 0)  # builds the plan the updates patch
        assert np.abs(r0.ranks - q).max() <= 1e-9
        for i, (b, y) in enumerate(seq):
            print(f"[update] blocking={blocking} batch={i} -{b.deletions}/+{b.insertions}", flush=True)
            ms = xg.update(b)
            assert ms > 0 and xg.M == y.M
            for which, off, idx in (("out", y.out_off, y.out_edg), ("in", y.in_off, y.in_edg)):
                go, gi = xg.csr(which)
                assert np.array_equal(go, off) and np.array_equal(gi, idx), (i, which)
            assert np.array_equal(c.mark_affected(xg, 2, b), oracle_mod.mark(2, y, b)), i
            for m in (0, 1, 2, 3, 4):
                res = solve(xg, m, batch=b, q=None if m == 0 else q)
                ref, it, _ = oracle_mod.pagerank_cuda(m, y, b, None if m == 0 else q)
                assert abs(res.iterations - it) <= 1, (i, m, res.iterations, it)
                assert np.abs(res.ranks - ref).max() <= 1e-9, (i, m)
            q, _, _ = oracle_mod.pagerank_cuda(0, y)
        xg.close()


def test_update_reports_time_and_keeps_the_plan(oracle_mod):
    """The first solve after an update reports the update in its preparation time; a blocked plan is not rebuilt
    (the next solve reports no plan build)."""
    from pagerank_amd import Batch, Context, HostGraph, Tuning, apply_batch, solve
    x = HostGraph.rmat(14, 16, seed=2)
    with Context(0) as c:
        c.set_tuning(Tuning(**{**c.get_tuning().__dict__, "blocking": 2048}))
        xg = c.upload(x)
        first = solve(xg, 0)
        assert first.preparationTime > 0  # layout and plan, built for this graph
        again = solve(xg, 0)
        assert again.preparationTime == 0
        b = Batch.random(x, 50, 0.8, seed=3)
        ms = xg.update(b)
        after = solve(xg, 4, batch=b, q=first.ranks)
        assert after.preparationTime >= ms * 0.999 and after.initializationTime >= after.preparationTime
        y = apply_batch(x, b)
        ref, it, _ = oracle_mod.pagerank_cuda(4, y, b, first.ranks)
        assert abs(after.iterations - it) <= 1 and np.abs(after.ranks - ref).max() <= 1e-9
        xg.close()


def test_update_errors_leave_graph_unchanged():
    from pagerank_amd import Batch, Context, HostGraph, PagerankError, engine
    x = HostGraph.rmat(10, 8, seed=5)
    with Context(0) as c:
        xg = c.upload(x)
        bad = Batch(np.array([0], np.uint32), np.array([x.N + 7], np.uint32), np.zeros(0, np.uint32),
                    np.zeros(0, np.uint32))
        with pytest.raises(PagerankError) as e:
            xg.update(bad)
        assert e.value.status == 1
        off, idx = xg.csr("out")
        assert np.array_equal(off, x.out_off) and np.array_equal(idx, x.out_edg)
        # after an update the batched and weighted solvers want the graph uploaded again
        xg.update(Batch.random(x, 20, 0.8, seed=1))
        with pytest.raises(PagerankError) as e:
            engine.ppr_batched(xg, [np.array([1, 2], np.uint32)])
        assert e.value.status == 5
        with pytest.raises(PagerankError) as e:
            engine.weighted_pagerank(xg, np.ones(xg.M, np.float32), np.full(x.N, 1.0 / x.N))
        assert e.value.status == 5
        static_only = c.upload(x, with_out=False)
        with pytest.raises(PagerankError) as e:
            static_only.update(Batch.random(x, 5, 0.8, seed=2))
        assert e.value.status == 5
        xg.close()
        static_only.close()


def test_deferred_in_csr(oracle_mod):
    """With the blocked plan patched, an update leaves the in-CSR unbuilt (the plan's base plus the net change since)
    and whatever reads it builds it: the CSR read-back after three updates with no reader between them (the net
    change composed over batches that re-insert deleted edges and delete inserted ones), the row plan, a rebuild of
    the blocked plan at another block size (the old base goes; the next update starts from the new plan), and a
    streaming session created from the updated graph. Blocked solves between the updates run on the unbuilt in-CSR.
    Bars as above: CSRs bit for bit, ranks within 1e-9 of the oracle, iterations within 1."""
    from pagerank_amd import Batch, Context, HostGraph, Stream, Tuning, apply_batch, solve
    x = HostGraph.rmat(13, 16, seed=12)
    seq = _series(x, 6, max(1, int(2e-3 * x.M)), 80)
    seq = seq[:2] + [seq[-1]] + seq[2:-1]  # the undo batch third, before any read-back

    def check(xg, y, tag):
        for which, off, idx in (("out", y.out_off, y.out_edg), ("in", y.in_off, y.in_edg)):
            go, gi = xg.csr(which)
            assert np.array_equal(go, off) and np.array_equal(gi, idx), (tag, which)

    def ranks_ok(xg, y, m, b, q, tag):
        res = solve(xg, m, batch=b, q=None if m == 0 else q)
        ref, it, _ = oracle_mod.pagerank_cuda(m, y, b, None if m == 0 else q)
        assert abs(res.iterations - it) <= 1, (tag, m, res.iterations, it)
        assert np.abs(res.ranks - ref).max() <= 1e-9, (tag, m)

    with Context(0) as c:
        base = c.get_tuning().__dict__
        c.set_tuning(Tuning(**{**base, "blocking": 1024}))
        xg = c.upload(x)
        q, _, _ = oracle_mod.pagerank_cuda(0, x)
        solve(xg, 0)  # builds the plan the updates patch
        # the undo batch undoes edges of the first two, so the first three are tidied against the graph they meet
        y = x
        for i in range(3):
            b = Batch.tidy(y, list(zip(seq[i][0].del_u.tolist(), seq[i][0].del_v.tolist())),
                           list(zip(seq[i][0].ins_u.tolist(), seq[i][0].ins_v.tolist())))
            print(f"[deferred] update {i}: -{b.deletions}/+{b.insertions}", flush=True)
            xg.update(b)
            y = apply_batch(y, b)
            ranks_ok(xg, y, 4, b, q, ("blocked", i))
            q, _, _ = oracle_mod.pagerank_cuda(0, y)
        check(xg, y, "three updates")
        for i in (3, 4):
            b = Batch.random(y, max(1, int(2e-3 * y.M)), 0.8, seed=90 + i)
            xg.update(b)
            y = apply_batch(y, b)
        print("[deferred] row plan", flush=True)
        c.set_tuning(Tuning(**{**base, "blocking": 0}))
        ranks_ok(xg, y, 0, None, None, "rows")
        ranks_ok(xg, y, 3, b, q, "rows")
        print("[deferred] blocked plan rebuilt at 2048", flush=True)
        c.set_tuning(Tuning(**{**base, "blocking": 2048}))
        ranks_ok(xg, y, 0, None, None, "rebuilt")
        q, _, _ = oracle_mod.pagerank_cuda(0, y)
        for i in (5, 6):
            b = Batch.random(y, max(1, int(2e-3 * y.M)), 0.8, seed=90 + i)
            xg.update(b)
            y = apply_batch(y, b)
            ranks_ok(xg, y, 4, b, q, ("rebuilt", i))
            q, _, _ = oracle_mod.pagerank_cuda(0, y)
        check(xg, y, "after the rebuild")
        print("[deferred] session from the updated graph", flush=True)
        b = Batch.random(y, 40, 0.8, seed=99)
        xg.update(b)
        y = apply_batch(y, b)
        st = Stream(xg, 1, max_batch=256)
        b2 = Batch.random(y, 40, 0.8, seed=100)
        st.update(b2, want_ranks=False)
        z = apply_batch(y, b2)
        for which, off, idx in (("out", z.out_off, z.out_edg), ("in", z.in_off, z.in_edg)):
            go, gi = st.csr(which)
            assert np.array_equal(go, off) and np.array_equal(gi, idx), ("session", which)
        st.close()
        c.set_tuning(Tuning(**base))
        xg.close()
