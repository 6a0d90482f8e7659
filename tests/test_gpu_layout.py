"""The two vertex orders of the layout every plan runs over (prh_tuning.layout): by in-degree (0, the default) and
the caller's order (1), switched back and forth on one graph.

Bars: every solver against the oracle (L-inf <= 1e-9, iterations within 1; DF-P against its own restatement) on
both orders and through both plans (row and propagation-blocked); the DT affected set bit-exact; a streaming session
in the caller's order equal to the sequential calls; the autotuner's choice on a grid against the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", ["grid", "rmat"])
@pytest.mark.parametrize("blocking", [0, 1024], ids=["rows", "blocked"])
def test_layouts_agree_with_oracle(oracle_mod, graph, blocking):
    from pagerank_amd import Batch, Context, HostGraph, Tuning, apply_batch, solve
    x = HostGraph.grid(70, 60, 0.62, seed=4) if graph == "grid" else HostGraph.rmat(12, 16, seed=21)
    b = Batch.random(x, 40, 0.8, seed=2)
    y = apply_batch(x, b)
    q, _, _ = oracle_mod.pagerank_cuda(0, x)
    refs = {m: oracle_mod.pagerank_cuda(m, y, b, None if m == 0 else q)[:2] for m in range(5)}
    with Context(0) as c:
        yg = c.upload(y)
        for layout in (1, 0, 1):
            print(f"[layout] graph={graph} blocking={blocking} layout={layout}", flush=True)
            c.set_tuning(Tuning(**{**c.get_tuning().__dict__, "blocking": blocking, "layout": layout}))
            for m in range(5):
                res = solve(yg, m, batch=b, q=None if m == 0 else q)
                ref, it = refs[m]
                assert abs(res.iterations - it) <= 1, (layout, m)
                assert np.abs(res.ranks - ref).max() <= 1e-9, (layout, m)
            assert np.array_equal(c.mark_affected(yg, 2, b), oracle_mod.mark(2, y, b))
        yg.close()


def test_stream_in_callers_order(oracle_mod):
    from pagerank_amd import Batch, Context, HostGraph, Stream, Tuning, apply_batch, solve
    x = HostGraph.grid(50, 40, 0.62, seed=7)
    q, _, _ = oracle_mod.pagerank_cuda(0, x)
    with Context(0) as c:
        c.set_tuning(Tuning(**{**c.get_tuning().__dict__, "blocking": 1024, "layout": 1}))
        xg = c.upload(x)
        st = Stream(xg, 4, max_batch=256, initial_ranks=q)
        y, prev = x, q
        for i in range(4):
            b = Batch.random(y, 30, 0.8, seed=70 + i)
            got = st.update(b)
            y = apply_batch(y, b)
            ref = solve(c.upload(y), 4, batch=b, q=prev)
            assert abs(got.iterations - ref.iterations) <= 1 and np.abs(got.ranks - ref.ranks).max() <= 1e-9, i
            off, idx = st.csr("out")
            assert np.array_equal(off, y.out_off) and np.array_equal(idx, y.out_edg), i
            prev = ref.ranks
        st.close()
        xg.close()


def test_autotune_on_grid(oracle_mod):
    """On a graph whose in-degrees stay at or below 64 the autotuner also times the caller's order; whatever it
    picks matches the oracle."""
    from pagerank_amd import Context, HostGraph, pagerankStaticHip
    y = HostGraph.grid(700, 600, 0.62, seed=3)  # 420K vertices, 1.5M edges
    ref, it, _ = oracle_mod.pagerank_cuda(0, y)
    with Context(0) as c:
        yg = c.upload(y, with_out=False)
        t, ms, _ = c.autotune(yg, None, iterations=3)
        print(f"[layout] autotune on grid 700x600: {t} at {ms:.4f} ms/iteration", flush=True)
        res = pagerankStaticHip(yg)
        yg.close()
    assert abs(res.iterations - it) <= 1 and np.abs(res.ranks - ref).max() <= 1e-9
