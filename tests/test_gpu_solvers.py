"""GPU parity of the five solvers, the degree partition and the affected marking, through the C ABI.

Bars (stated per assertion):
  * ranks vs the oracle's restatement of the reference GPU path: L-inf <= 1e-9 (FP64; only summation order differs);
  * iterations vs the oracle: equal or +-1 (the L-inf change can land within rounding of the tolerance);
  * Static / ND / DT / DF vs the compensated FP64 power iteration (the exact fixed point): L-inf <= 1e-9;
  * DF-P vs the power iteration: the pruning method is approximate by design (vertices whose relative change
    is <= 1e-6 stop updating), so its own bound is 1e-5 relative to the largest rank;
  * degree partition and affected flags: bit-exact.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import complete_graph, cycle_graph, star_graph, torus_graph

pytestmark = pytest.mark.gpu

RANK_TOL = 1e-9


def _graphs():
    from pagerank_amd import HostGraph
    return {
        "rmat12": lambda: HostGraph.rmat(12, 16, seed=7),
        "rmat14": lambda: HostGraph.rmat(14, 16, seed=11),
        "grid60x50": lambda: HostGraph.grid(60, 50, 0.62, seed=5),
        "star5000": lambda: star_graph(5000),
    }


GRAPH_NAMES = ["rmat12", "rmat14", "grid60x50", "star5000"]


@pytest.fixture(scope="module")
def graphs():
    return {k: f() for k, f in _graphs().items()}


def _close_iters(a: int, b: int) -> bool:
    return abs(a - b) <= 1


@pytest.mark.parametrize("name", GRAPH_NAMES)
def test_static_matches_oracle_and_fixed_point(ctx, oracle_mod, graphs, name):
    from pagerank_amd import pagerankStaticHip
    y = graphs[name]
    print(f"[static] graph={name} N={y.N} M={y.M}", flush=True)
    yg = ctx.upload(y)
    res = pagerankStaticHip(yg)
    ref, it, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y)
    exact, _ = oracle_mod.power_iteration(y)
    assert _close_iters(res.iterations, it), (res.iterations, it)
    assert np.abs(res.ranks - ref).max() <= RANK_TOL
    assert np.abs(res.ranks - exact).max() <= RANK_TOL
    assert abs(res.ranks.sum() - 1.0) < 1e-9  # no dead ends: ranks sum to 1
    assert res.ranks.min() >= 0.15 / y.N * (1 - 1e-9)  # r_v >= (1-alpha)/N
    assert res.computationTime > 0 and res.time >= res.computationTime


@pytest.mark.parametrize("make", [lambda: cycle_graph(1000), lambda: complete_graph(70), lambda: torus_graph(40, 30)],
                         ids=["cycle", "complete", "torus"])
def test_vertex_transitive_closed_form(ctx, make):
    """Vertex-transitive graphs with self-loops have PageRank exactly 1/N; from r0 = 1/N one iteration suffices."""
    from pagerank_amd import pagerankStaticHip
    y = make()
    res = pagerankStaticHip(ctx.upload(y))
    assert res.iterations == 1
    assert np.abs(res.ranks - 1.0 / y.N).max() <= 1e-15


def test_dense_solve_small(ctx):
    """(I - alpha A) r = (1-alpha)/N 1, solved densely (numpy.linalg.solve), pins the fixed point."""
    from pagerank_amd import HostGraph, PagerankOptions, pagerankStaticHip
    y = HostGraph.rmat(8, 8, seed=3)
    A = y.dense_transition()
    exact = np.linalg.solve(np.eye(y.N) - 0.85 * A, np.full(y.N, 0.15 / y.N))
    res = pagerankStaticHip(ctx.upload(y), PagerankOptions(tolerance=1e-14))
    assert np.abs(res.ranks - exact).max() <= 1e-12


@pytest.mark.parametrize("frac", [1e-3, 1e-2])
@pytest.mark.parametrize("method", [1, 2, 3, 4], ids=["ND", "DT", "DF", "DFP"])
def test_dynamic_matches_oracle(ctx, oracle_mod, method, frac):
    from pagerank_amd import Batch, HostGraph, apply_batch, solve
    x = HostGraph.rmat(13, 16, seed=21)
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    b = Batch.random(x, max(1, int(frac * x.M)), 0.8, seed=9)
    y = apply_batch(x, b)
    print(f"[dynamic] method={method} frac={frac} N={y.N} M={y.M} del={b.deletions} ins={b.insertions}", flush=True)
    res = solve(ctx.upload(y), method, batch=b, q=q)
    ref, it, _ = oracle_mod.pagerank_cuda(method, y, b, q)
    exact, _ = oracle_mod.power_iteration(y)
    assert _close_iters(res.iterations, it), (res.iterations, it)
    assert np.abs(res.ranks - ref).max() <= RANK_TOL
    if method == 4:
        # Pruned vertices stop updating while their neighbours still move, so DF-P stops short of the exact
        # fixed point (about 3.5e-6 of the largest rank on this graph, in the oracle and on the GPU alike).
        assert np.abs(res.ranks - exact).max() <= 1e-5 * exact.max()
        assert np.abs(ref - exact).max() <= 1e-5 * exact.max()
    else:
        assert np.abs(res.ranks - exact).max() <= RANK_TOL


# (switch_degree, huge_degree, low_width, source_slices): row plans (slices 0) and source-sliced plans.
PLANS = [(128, 16384, 0, 0), (8, 64, 0, 0), (16, 100, 4, 0), (2, 2, 1, 0), (32, 33, 32, 0), (64, 1000, 2, 0),
         (128, 16384, 0, 8), (64, 64, 0, 8), (128, 200, 0, 4), (2, 2, 0, 2), (128, 16384, 0, 1), (64, 64, 0, 1)]


@pytest.mark.parametrize("plan", PLANS, ids=[f"sw{a}_huge{b}_w{c}_sl{d}" for a, b, c, d in PLANS])
@pytest.mark.parametrize("method", [0, 1, 4], ids=["static", "ND", "DFP"])
def test_launch_plans_agree(ctx, oracle_mod, plan, method):
    """Every launch plan (chunked rows with their second combining launch, wave rows, fixed or banded lanes per
    row; for Static and ND also the source-sliced gather with its per-row slice combine) gives the oracle's ranks
    to 1e-9: the plan only changes the summation order."""
    from pagerank_amd import Batch, HostGraph, Tuning, apply_batch, solve
    x = HostGraph.rmat(12, 24, seed=4)
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    b = Batch.random(x, 200, 0.8, seed=3)
    y = apply_batch(x, b)
    old = ctx.get_tuning()
    print(f"[plan] {plan} method={method} N={y.N} M={y.M} max_in={int(np.diff(y.in_off).max())}", flush=True)
    try:
        ctx.set_tuning(Tuning(plan[0], plan[1], plan[2], old.check_interval, plan[3]))
        res = solve(ctx.upload(y), method, batch=b if method else None, q=q if method else None)
    finally:
        ctx.set_tuning(old)
    ref, it, _ = oracle_mod.pagerank_cuda(method, y, b if method else None, q if method else None)
    assert _close_iters(res.iterations, it), (res.iterations, it)
    assert np.abs(res.ranks - ref).max() <= RANK_TOL


@pytest.mark.parametrize("slices", [0, 1, 8])
def test_rows_without_in_edges(ctx, oracle_mod, slices):
    """Vertices nobody links to (no self-loops here) keep rank (1-alpha)/N; a row with no sources has no item in
    any slice and its sum must be exactly 0."""
    from pagerank_amd import HostGraph, Tuning, pagerankStaticHip
    n = 3000
    u = np.arange(n, dtype=np.uint32)
    src = np.concatenate([u[1:], [0], u[:100]])
    dst = np.concatenate([np.zeros(n - 1, np.uint32), [1], (u[:100] * 7 + 3) % 50])
    y = HostGraph.from_edges(n, src, dst, self_loops=False)
    assert (np.diff(y.in_off.astype(np.int64)) == 0).sum() > n // 2
    old = ctx.get_tuning()
    print(f"[no-in-edges] slices={slices} N={y.N} M={y.M}", flush=True)
    try:
        ctx.set_tuning(Tuning(64, 64, old.low_width, old.check_interval, slices))
        res = pagerankStaticHip(ctx.upload(y))
    finally:
        ctx.set_tuning(old)
    ref, it, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y)
    assert _close_iters(res.iterations, it)
    assert np.abs(res.ranks - ref).max() <= RANK_TOL
    zero_in = np.diff(y.in_off.astype(np.int64)) == 0
    assert np.all(res.ranks[zero_in] == (1 - 0.85) / n)  # C0 exactly: fma(alpha, 0, C0)


@pytest.mark.parametrize("method", [2, 3, 4], ids=["DT", "DF", "DFP"])
def test_empty_batch_returns_q(ctx, oracle_mod, method):
    """With nothing affected, one iteration copies q through (L-inf change 0 < tolerance)."""
    from pagerank_amd import Batch, HostGraph, solve
    y = HostGraph.rmat(10, 16, seed=2)
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y)
    res = solve(ctx.upload(y), method, batch=Batch.empty(), q=q)
    assert res.iterations == 1
    assert np.array_equal(res.ranks, q)


def test_warm_start_from_solution_converges_in_one(ctx, oracle_mod):
    from pagerank_amd import HostGraph, PagerankOptions, pagerankNaiveDynamicHip
    y = HostGraph.rmat(11, 16, seed=4)
    exact, _ = oracle_mod.power_iteration(y)
    res = pagerankNaiveDynamicHip(ctx.upload(y), exact, PagerankOptions())
    assert res.iterations == 1


def test_resume_is_bit_exact(ctx):
    """Static with k iterations, then ND from that result for the rest, equals the uninterrupted run bit for bit
    (a Jacobi step is a pure function of the rank vector)."""
    from pagerank_amd import HostGraph, PagerankOptions, pagerankNaiveDynamicHip, pagerankStaticHip
    y = HostGraph.rmat(12, 16, seed=8)
    yg = ctx.upload(y)
    full = pagerankStaticHip(yg, PagerankOptions(tolerance=0.0, maxIterations=40))
    part = pagerankStaticHip(yg, PagerankOptions(tolerance=0.0, maxIterations=15))
    rest = pagerankNaiveDynamicHip(yg, part.ranks, PagerankOptions(tolerance=0.0, maxIterations=25))
    assert full.iterations == 40 and part.iterations == 15 and rest.iterations == 25
    assert np.array_equal(full.ranks, rest.ranks)


@pytest.mark.parametrize("n", [1, 2, 63, 2047, 2048, 2049, 100_003])
def test_partition_bit_exact(ctx, oracle_mod, n):
    rng = np.random.default_rng(n)
    deg = rng.choice([0, 1, 63, 64, 65, 200, 5000], size=n, p=[.1, .3, .1, .1, .1, .2, .1]).astype(np.int64)
    off = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(deg, out=off[1:])
    order, npart = ctx.partition(off, 64)
    ref_order, ref_np = oracle_mod.partition(off, 64)
    assert npart == ref_np
    assert np.array_equal(order, ref_order)


@pytest.mark.parametrize("method", [2, 3], ids=["DT", "DF"])
@pytest.mark.parametrize("frac", [1e-4, 1e-3, 5e-2])
def test_affected_marking_bit_exact(ctx, oracle_mod, method, frac):
    from pagerank_amd import Batch, HostGraph, apply_batch
    x = HostGraph.rmat(13, 8, seed=33)
    b = Batch.random(x, max(1, int(frac * x.M)), 0.8, seed=17)
    y = apply_batch(x, b)
    got = ctx.mark_affected(ctx.upload(y), method, b)
    want = oracle_mod.mark(method, y, b)
    assert np.array_equal(got, want)
    if oracle_mod.ref_available():  # the reference's CPU marking gives the same set
        assert np.array_equal(got, oracle_mod.ref_affected(0 if method == 3 else 1, x, b))


def test_grid_traversal_long_chain(ctx, oracle_mod):
    """A grid has a large diameter: the traversal closure needs many breadth-first levels."""
    from pagerank_amd import Batch, HostGraph, apply_batch
    x = HostGraph.grid(40, 40, 0.9, seed=1)
    b = Batch.tidy(x, [[0, 1]], [[5, 700]])
    y = apply_batch(x, b)
    got = ctx.mark_affected(ctx.upload(y), 2, b)
    assert np.array_equal(got, oracle_mod.mark(2, y, b))


@pytest.mark.parametrize("frac", [1e-4, 1e-2])
@pytest.mark.parametrize("method", [3, 4], ids=["DF", "DFP"])
def test_extension_directions_agree(ctx, oracle_mod, method, frac):
    """Push, pull and the per-iteration choice between them set the same affected flags, so DF and DF-P give
    bit-identical ranks and iteration counts with every extension_direction, and the oracle's ranks to 1e-9."""
    from pagerank_amd import Batch, HostGraph, Tuning, apply_batch, solve
    x = HostGraph.rmat(13, 16, seed=8)
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    b = Batch.random(x, max(1, int(frac * x.M)), 0.8, seed=12)
    y = apply_batch(x, b)
    yg = ctx.upload(y)
    old = ctx.get_tuning()
    out = {}
    try:
        for d in (0, 1, 2):
            print(f"[direction] method={method} frac={frac} direction={d}", flush=True)
            t = ctx.get_tuning()
            t.extension_direction = d
            ctx.set_tuning(t)
            out[d] = solve(yg, method, batch=b, q=q)
    finally:
        ctx.set_tuning(old)
    ref, it, _ = oracle_mod.pagerank_cuda(method, y, b, q)
    for d in (1, 2):
        assert out[d].iterations == out[0].iterations
        assert np.array_equal(out[d].ranks, out[0].ranks), d
    assert _close_iters(out[0].iterations, it)
    assert np.abs(out[0].ranks - ref).max() <= RANK_TOL


def test_hub_source_chunks_and_schedule_kept_across_tunings(ctx, oracle_mod):
    """A source with 6000 out-edges is cut into chunks by the extension's cached out-degree schedule (rows above
    4096 out-edges). DT marking and DF through the row plan match the oracle, and a tuning change between two
    solves, which rebuilds the rank plan, leaves the graph's cached schedule intact."""
    from pagerank_amd import Batch, HostGraph, apply_batch, solve
    n = 9000
    rng = np.random.default_rng(4)
    src = np.concatenate([np.zeros(6000, np.uint32), rng.integers(0, n, 30000).astype(np.uint32)])
    dst = np.concatenate([np.arange(1, 6001, dtype=np.uint32), rng.integers(0, n, 30000).astype(np.uint32)])
    x = HostGraph.from_edges(n, src, dst, self_loops=True)
    b = Batch.tidy(x, [[0, 1]], [[0, 7000], [5, 8000]])  # the hub is a batch source
    y = apply_batch(x, b)
    print(f"[hub] N={y.N} M={y.M} hub out-degree={int(y.out_off[1] - y.out_off[0])}", flush=True)
    yg = ctx.upload(y)
    assert np.array_equal(ctx.mark_affected(yg, 2, b), oracle_mod.mark(2, y, b))
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    ref, it, _ = oracle_mod.pagerank_cuda(3, y, b, q)
    old = ctx.get_tuning()
    try:
        first = solve(yg, 3, batch=b, q=q)
        t = ctx.get_tuning()
        t.switch_degree = 128
        ctx.set_tuning(t)
        second = solve(yg, 3, batch=b, q=q)
    finally:
        ctx.set_tuning(old)
    assert _close_iters(first.iterations, it) and _close_iters(second.iterations, it)
    assert np.abs(first.ranks - ref).max() <= RANK_TOL
    assert np.abs(second.ranks - ref).max() <= RANK_TOL
