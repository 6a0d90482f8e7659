"""The banded plan (prh_tuning.window = 1, rows in the caller's order) against the oracle, through the C ABI.

Each workgroup sweeps a contiguous range of rows chunk by chunk (1024 rows) through a ring of contributions in LDS
that holds the window [c0 - H, c0 + 1024 + H) around the chunk; sources outside it are gathered from the ranks. The
graphs cover: grids whose vertical edges span H = cols (the smaller ring, two workgroups per CU, at H = 800; the
larger ring, one per CU, at H = 9000), several chunks per workgroup, a grid whose spread (12000) exceeds what the
larger ring holds (its vertical edges are gathered outside the window), RMAT (no locality, rows with more in-edges
than a row loads ahead of its chunk), a graph smaller than one chunk, and random insertions that leave the band.

Bars (the row plan's): ranks within L-inf 1e-9 of the oracle's restatement of the reference GPU path, iterations
equal or +-1; every sum is taken in a fixed order, so two solves agree bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graph(name):
    from pagerank_amd import HostGraph
    if name == "grid1500x800":
        return HostGraph.grid(1500, 800, 0.62, seed=5)
    if name == "grid150x9000":
        return HostGraph.grid(150, 9000, 0.62, seed=6)
    if name == "grid100x12000":
        return HostGraph.grid(100, 12000, 0.62, seed=7)
    if name == "grid7x5":
        return HostGraph.grid(7, 5, 0.62, seed=8)
    if name == "rmat14":
        return HostGraph.rmat(14, 16, seed=11)
    raise KeyError(name)


def _banded(ctx, **kw):
    from pagerank_amd import Tuning
    ctx.set_tuning(Tuning(**{**ctx.get_tuning().__dict__, "blocking": 0, "layout": 1, "window": 1, **kw}))


GRAPHS = ["grid1500x800", "grid150x9000", "grid100x12000", "grid7x5", "rmat14"]


@pytest.mark.parametrize("name", GRAPHS)
def test_window_static_matches_oracle(oracle_mod, name):
    from pagerank_amd import Context, pagerankStaticHip
    y = _graph(name)
    ref, it, _ = oracle_mod.pagerank_cuda(0, y)
    with Context(0) as c:
        _banded(c)
        yg = c.upload(y, with_out=False)
        print(f"[window] {name}: N={y.N} M={y.M}", flush=True)
        res = pagerankStaticHip(yg)
        again = pagerankStaticHip(yg)
        yg.close()
    err = float(np.abs(res.ranks - ref).max())
    print(f"[window] {name}: {res.iterations} iterations (oracle {it}), L-inf {err:.3e}", flush=True)
    assert abs(res.iterations - it) <= 1
    assert err <= 1e-9
    assert np.array_equal(res.ranks, again.ranks)  # fixed summation order: repeats bit for bit


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_window_dynamic_methods(oracle_mod, direction):
    """ND / DT / DF / DF-P after a batch whose random insertions leave the band (gathered outside the window)."""
    from pagerank_amd import Batch, Context, apply_batch, solve
    x = _graph("grid1500x800")
    b = Batch.random(x, 2000, 0.8, seed=3)
    y = apply_batch(x, b)
    q, _, _ = oracle_mod.pagerank_cuda(0, x)
    with Context(0) as c:
        _banded(c, extension_direction=direction)
        yg = c.upload(y)
        for m in range(5):
            print(f"[window] dynamic method {m}, extension direction {direction}", flush=True)
            ref, it, _ = oracle_mod.pagerank_cuda(m, y, b, None if m == 0 else q)
            res = solve(yg, m, batch=b, q=None if m == 0 else q)
            assert abs(res.iterations - it) <= 1, m
            assert np.abs(res.ranks - ref).max() <= 1e-9, m
        yg.close()


def test_window_after_inplace_update(oracle_mod):
    """A graph updated on the device keeps its banded plan; the inserted edges outside the window are gathered."""
    from pagerank_amd import Batch, Context, apply_batch, solve
    x = _graph("grid150x9000")
    q, _, _ = oracle_mod.pagerank_cuda(0, x)
    with Context(0) as c:
        _banded(c)
        xg = c.upload(x)
        solve(xg, 0)
        y = x
        for i in range(2):
            b = Batch.random(y, 500, 0.8, seed=40 + i)
            xg.update(b)
            y = apply_batch(y, b)
            print(f"[window] update {i}: -{b.deletions}/+{b.insertions}", flush=True)
            for m in (0, 4):
                ref, it, _ = oracle_mod.pagerank_cuda(m, y, b, None if m == 0 else q)
                res = solve(xg, m, batch=b, q=None if m == 0 else q)
                assert abs(res.iterations - it) <= 1, (i, m)
                assert np.abs(res.ranks - ref).max() <= 1e-9, (i, m)
            q = oracle_mod.pagerank_cuda(0, y)[0]
        xg.close()


def test_window_needs_callers_order():
    from pagerank_amd import Context, PagerankError, Tuning
    with Context(0) as c:
        with pytest.raises(PagerankError):
            c.set_tuning(Tuning(**{**c.get_tuning().__dict__, "layout": 0, "window": 1}))
