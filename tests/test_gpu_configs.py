"""Every benchmark configuration at its stated size on the GPU, checked against the pinned CPU oracle.

The reference checks each solver on every batch against a converged CPU solve (main.cu:55, :65, :88). Here:
  C2a  static PageRank on directed RMAT scale 24 (edge factor 16; 16.8M vertices, 280M edges), autotuned plan
       (the propagation-blocked update with B = 8192 on this graph: 2048 bins, most split across items);
  C2b  static PageRank on the 6000 x 4000 road-like grid (24M vertices, 83.5M edges), autotuned plan;
  C4   64-column batched personalized PageRank on RMAT scale 22 (4.2M vertices, 69M edges), and 16 DF-P batches of
       1e-4 |E| through the streaming session;
  C5   weighted PageRank and its adjoint on RMAT scale 20 (1M vertices, 17M edges) with FP32 weights U(0.5, 1.5),
       and central finite differences on 32 edges.
Static results are compared with the oracle's OpenMP port of pagerankStaticOmp (bit-identical to the compiled
reference on the test graphs); PPR columns with the oracle's single-column solve; the weighted solves with the
oracle's forward, adjoint and gradient. Bars: L-inf <= 1e-9, iterations within 1, ranks summing to 1 and no rank
below (1 - alpha) / N (no dead ends); DF-P within 1e-5 of the largest rank (its pruning is approximate).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALPHA = 0.85


def _invariants(r: np.ndarray):
    n = len(r)
    assert abs(r.sum() - 1.0) <= 1e-9
    assert r.min() >= (1 - ALPHA) / n * (1 - 1e-9)


def _static_against_port(graph, oracle_mod, say, label: str):
    from pagerank_amd import Context, pagerankStaticHip
    with Context(0) as c:
        say(f"{label}: N={graph.N} M={graph.M}: upload and autotune")
        yg = c.upload(graph, with_out=False)
        t, ms, _ = c.autotune(yg, None, iterations=5)
        res = pagerankStaticHip(yg)
        yg.close()
        say(f"{label}: GPU {res.iterations} iterations, {ms:.3f} ms/iteration with blocking={t.blocking}; "
            f"oracle port ...")
    ref, it, cpu_ms = oracle_mod.static_omp(graph)
    err = float(np.abs(res.ranks - ref).max())
    say(f"{label}: port {it} iterations in {cpu_ms / 1e3:.1f} s; L-inf GPU vs port {err:.3e}")
    assert abs(res.iterations - it) <= 1
    assert err <= 1e-9
    _invariants(res.ranks)
    return t


def test_c2a_rmat24_static(oracle_mod, say):
    from pagerank_amd import HostGraph
    g = HostGraph.rmat(24, 16, seed=42)
    t = _static_against_port(g, oracle_mod, say, "C2a RMAT-24")
    assert t.blocking >= 4096  # the blocked plan with many bins is the one the autotuner runs at this size


def test_c2b_grid_static(oracle_mod, say):
    from pagerank_amd import HostGraph
    g = HostGraph.grid(6000, 4000, 0.62, seed=42)
    _static_against_port(g, oracle_mod, say, "C2b grid 6000x4000")


@pytest.fixture(scope="module")
def rmat22():
    from pagerank_amd import HostGraph
    return HostGraph.rmat(22, 16, seed=42)


def test_c4_batched_ppr_64_columns(rmat22, oracle_mod, say):
    """64 seed sets of 1..1000 vertices (column 63: every vertex, whose solve is the static one)."""
    from pagerank_amd import Context, engine, pagerankStaticHip
    from pagerank_amd.experiment import seed_sets
    y = rmat22
    rng = np.random.default_rng(42)
    sets = seed_sets(y.N, 64, rng)
    sets[63] = np.arange(y.N, dtype=np.uint32)
    with Context(0) as c:
        yg = c.upload(y, with_out=False)
        say(f"C4 RMAT-22: N={y.N} M={y.M}: 64-column batched solve")
        res = engine.ppr_batched(yg, sets)
        static = pagerankStaticHip(yg)
        yg.close()
    assert res.ranks.shape == (y.N, 64)
    sums = res.ranks.sum(axis=0)
    assert np.abs(sums - 1.0).max() <= 1e-9
    assert abs(int(res.iterations[63]) - static.iterations) <= 1
    assert np.abs(res.ranks[:, 63] - static.ranks).max() <= 1e-9
    for j in (0, 21, 42, 62):
        s = np.zeros(y.N)
        np.add.at(s, sets[j], 1.0 / len(sets[j]))
        ref, it = oracle_mod.ppr(y, s)
        err = float(np.abs(res.ranks[:, j] - ref).max())
        say(f"C4 column {j} (|seeds|={len(sets[j])}): {res.iterations[j]} iterations (oracle {it}), L-inf {err:.3e}")
        assert abs(int(res.iterations[j]) - it) <= 1
        assert err <= 1e-9


def test_c4_stream_16_dfp_batches(rmat22, oracle_mod, say):
    """16 consecutive DF-P batches of 1e-4 |E| (80% insertions) through the streaming session on the autotuned plan
    (the propagation-blocked update, patched per batch); after every batch the session's out-CSR equals the
    host-applied graph's bit for bit; the last ranks against the port's converged solve of that graph."""
    from pagerank_amd import PRUNE_DYNAMIC_FRONTIER, Batch, Context, Stream, apply_batch
    x = rmat22
    size = int(1e-4 * x.M)
    with Context(0) as c:
        xg = c.upload(x)
        t, _, _ = c.autotune(xg, None, iterations=3)
        st = Stream(xg, PRUNE_DYNAMIC_FRONTIER, max_batch=size + 16)
        y = x
        for i in range(16):
            b = Batch.random(y, size, 0.8, seed=500 + i)
            got = st.update(b, want_ranks=False)
            y = apply_batch(y, b)
            assert st.size == y.M, i
            off, idx = st.csr("out")
            assert np.array_equal(off, y.out_off) and np.array_equal(idx, y.out_edg), i
            if i % 4 == 3:
                say(f"C4b batch {i} (blocking {t.blocking}): -{b.deletions}/+{b.insertions}, {got.iterations} "
                    f"iterations, {got.time:.2f} ms")
        last = st.ranks()
        st.close()
        xg.close()
    ref, _, _ = oracle_mod.static_omp(y)
    err = float(np.abs(last - ref).max())
    say(f"C4b after 16 batches: L-inf vs the port's solve of the host-applied graph {err:.3e}")
    assert err <= 1e-5 * ref.max()


def test_c5_rmat20_adjoint(oracle_mod, say):
    """Forward, adjoint and both gradients against the oracle; dL/dw against central differences on 32 edges."""
    torch = pytest.importorskip("torch")
    from pagerank_amd import Context, HostGraph, PagerankOptions, Tuning, weighted_adjoint, weighted_pagerank
    from pagerank_amd.autograd import pagerank
    y = HostGraph.rmat(20, 16, seed=42)
    rng = np.random.default_rng(20)
    w = rng.uniform(0.5, 1.5, y.M).astype(np.float32)
    cv = rng.standard_normal(y.N)
    s = np.full(y.N, 1.0 / y.N)
    with Context(0) as c:
        yg = c.upload(y)
        t, _, _ = c.autotune(yg, None, iterations=5)
        say(f"C5 RMAT-20: N={y.N} M={y.M}, blocking={t.blocking}: forward and adjoint")
        fw = weighted_pagerank(yg, w, s)
        g = 2 * cv * fw.ranks
        lam, gw, gs, ait = weighted_adjoint(yg, w, fw.ranks, g)
        r, it, W = oracle_mod.weighted_forward(y, w, s)
        assert abs(fw.iterations - it) <= 1
        assert np.abs(fw.ranks - r).max() <= 1e-9
        lam_o, ait_o = oracle_mod.weighted_adjoint(y, w, W, g)
        gw_o, gs_o = oracle_mod.weighted_grad(y, W, fw.ranks, lam_o, g)
        assert abs(ait - ait_o) <= 1
        for got, want in ((lam, lam_o), (gw, gw_o), (gs, gs_o)):
            assert np.abs(got - want).max() <= 1e-9
        # Central differences through the autograd harness, forward solves run to stagnation. The loss is scaled by
        # N^2 so that ranks of order 1/N give gradients of order one; the bar is relative to the largest gradient.
        # One edge moves L by about 1e-14 of its size here, below the rounding of L itself, so the difference is
        # taken per vertex: L(w+) - L(w-) = sum_v c_v (r+_v - r-_v)(r+_v + r-_v).
        dev = torch.device("cuda:0")
        scale = float(y.N) ** 2
        cw = torch.tensor(cv * scale, device=dev)
        wt = torch.tensor(w, device=dev, requires_grad=True)
        st = torch.tensor(s, device=dev, requires_grad=True)
        (cw * pagerank(wt, st, yg, PagerankOptions()) ** 2).sum().backward()
        grad = wt.grad.double().cpu().numpy()
        stag = PagerankOptions(tolerance=0.0, maxIterations=400)
        # The differences use the row plan: its sums are taken in a fixed order, so the two perturbed solves round
        # alike and their difference is not swamped by the blocked plan's order-dependent rounding (about 1e-16 of
        # every rank, against a change of about 1e-14 of L from one edge).
        c.set_tuning(Tuning(**{**t.__dict__, "blocking": 0}))

        def R(wv):
            with torch.no_grad():
                return pagerank(torch.tensor(wv, device=dev), st.detach(), yg, stag)

        edges = np.random.default_rng(3).choice(y.M, size=32, replace=False)
        gmax = float(np.abs(grad[edges]).max())
        # The difference cannot resolve less than the rounding of the two solves: a rank within half an ulp
        # (2^-53 relative) of the other solve's moves c_v r_v^2 by up to 2^-52 of itself, so the two losses differ
        # by up to 2 * 2^-52 * sum_v |c_v| r_v^2 before any change of w. On RMAT-20 that floor is about 7e-4 / h
        # against gradients up to about 1e2 (its rms over the vertices is about 1e-4); a dropped or misplaced term
        # in the gradient moves it by a sizeable part of itself, far above the floor.
        terms = float((np.abs(cv) * scale * fw.ranks * fw.ranks).sum())
        worst = 0.0
        for e in edges:
            wp, wm = w.copy(), w.copy()
            wp[e] = np.float32(w[e] + np.float32(1e-6))
            wm[e] = np.float32(w[e] - np.float32(1e-6))
            rp, rm = R(wp), R(wm)
            h2 = float(wp[e]) - float(wm[e])
            fd = float((cw * (rp - rm) * (rp + rm)).sum()) / h2
            floor = 2.0 * 2.0 ** -52 * terms / h2
            worst = max(worst, abs(fd - grad[e]) / (1e-4 * abs(fd) + floor))
            assert abs(fd - grad[e]) <= 1e-4 * abs(fd) + floor, (e, fd, grad[e], floor)
        say(f"C5 finite differences on 32 edges (h = 1e-6, gradients up to {gmax:.1f}, rounding floor "
            f"{2.0 * 2.0 ** -52 * terms / 2e-6:.1e}): worst error / bar = {worst:.2f}")
        yg.close()
