"""Pins of the oracle (CPU, no GPU needed).

The oracle (oracle/prh_oracle.c) is checked against
  * the reference's own CPU path, through committed fixtures (tests/golden/*.npz, written by make_golden.py from
    oracle/_ref) and, where oracle/_ref is built, live on fresh graphs;
  * closed forms: vertex-transitive graphs with self-loops have PageRank exactly 1/N;
  * a library routine: the dense solve (I - alpha A) r = (1-alpha)/N 1 (numpy.linalg.solve);
  * invariants: ranks sum to 1 with no dead ends, r_v >= (1-alpha)/N;
  * brute force: the stable degree partition against a direct Python statement, the affected sets against a
    Python breadth-first search.
The last test plants plausible mistakes in the oracle's inputs (a transposed operand, a wrong damping term, an
off-by-one source index) and shows the pins reject them.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from conftest import complete_graph, cycle_graph, star_graph, torus_graph

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def _load(name):
    from pagerank_amd import Batch, HostGraph
    z = np.load(GOLDEN / f"{name}.npz")  # data only (allow_pickle stays False)
    x = HostGraph.from_csr(z["x_off"], z["x_edg"])
    b = Batch(z["del_u"], z["del_v"], z["ins_u"], z["ins_v"])
    return z, x, b


def test_fixtures_present():
    assert set(FIXTURES) >= {"rmat8", "rmat10", "grid24"}


@pytest.mark.parametrize("name", FIXTURES)
def test_batch_application_matches_reference(name):
    from pagerank_amd import apply_batch
    z, x, b = _load(name)
    y = apply_batch(x, b)
    assert np.array_equal(y.out_off, z["y_off"]) and np.array_equal(y.out_edg, z["y_edg"])


@pytest.mark.parametrize("name", FIXTURES)
def test_static_port_bit_exact_with_reference(oracle_mod, name):
    """The OpenMP port (the CPU baseline) reproduces pagerankStaticOmp bit for bit."""
    from pagerank_amd import apply_batch
    z, x, b = _load(name)
    y = apply_batch(x, b)
    r, it, _ = oracle_mod.static_omp(y)
    assert it == int(z["static_iterations"])
    assert np.array_equal(r, z["static_ranks"])


ASYNC = [(0, "static"), (1, "ND"), (2, "DT"), (3, "DF"), (4, "DFP")]


@pytest.mark.parametrize("method,tag", ASYNC, ids=[t for _, t in ASYNC])
@pytest.mark.parametrize("name", FIXTURES)
def test_async_port_bit_exact_with_reference(oracle_mod, name, method, tag):
    """On one thread the asynchronous OpenMP solvers (the dynamic CPU baselines of the reference's experiment)
    sweep in vertex order, and the port reproduces the reference's ranks and iteration counts bit for bit."""
    from pagerank_amd import apply_batch
    z, x, b = _load(name)
    y = apply_batch(x, b)
    r, it, _ = oracle_mod.pagerank_async(method, y, b if method else None, z["q"] if method else None, threads=1)
    assert it == int(z[f"async_{tag}_iterations"])
    assert np.array_equal(r, z[f"async_{tag}_ranks"])


@pytest.mark.parametrize("method", [0, 1, 2, 3, 4])
def test_async_port_reaches_fixed_point(oracle_mod, method):
    """With several threads the in-place reads race, as in the reference, but every sweep is still a contraction:
    Static/ND/DT/DF end within 1e-8 of the exact fixed point, DF-P within its pruning bound (1e-5 x max rank)."""
    from pagerank_amd import Batch, HostGraph, apply_batch
    x = HostGraph.rmat(11, 16, seed=17)
    q, _, _ = oracle_mod.static_omp(x)
    b = Batch.random(x, 200, 0.8, seed=8)
    y = apply_batch(x, b)
    exact, _ = oracle_mod.power_iteration(y)
    r, it, _ = oracle_mod.pagerank_async(method, y, b if method else None, q if method else None, threads=4)
    bound = 1e-5 * exact.max() if method == 4 else 1e-8
    assert 0 < it < 500 and np.abs(r - exact).max() <= bound


@pytest.mark.parametrize("name", FIXTURES)
def test_cuda_restatement_static_nd(oracle_mod, name):
    """The GPU-path restatement differs from the CPU path only by fused multiply-adds and the 256-way block sum."""
    from pagerank_amd import apply_batch
    z, x, b = _load(name)
    y = apply_batch(x, b)
    r, it, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y)
    assert abs(it - int(z["static_iterations"])) <= 1
    assert np.abs(r - z["static_ranks"]).max() <= 1e-15
    r, it, _ = oracle_mod.pagerank_cuda(oracle_mod.NAIVE_DYNAMIC, y, b, z["q"])
    assert abs(it - int(z["ND_iterations"])) <= 1
    assert np.abs(r - z["ND_ranks"]).max() <= 1e-15


@pytest.mark.parametrize("name", FIXTURES)
def test_cuda_restatement_dynamic_traversal(oracle_mod, name):
    """DT updates a fixed affected set synchronously, as the CPU path does: same iterations, same ranks."""
    from pagerank_amd import apply_batch
    z, x, b = _load(name)
    y = apply_batch(x, b)
    r, it, vaff = oracle_mod.pagerank_cuda(oracle_mod.DYNAMIC_TRAVERSAL, y, b, z["q"])
    assert np.array_equal(vaff, z["affected_dt"])
    assert abs(it - int(z["DT_iterations"])) <= 1
    assert np.abs(r - z["DT_ranks"]).max() <= 1e-15


@pytest.mark.parametrize("name", FIXTURES)
def test_affected_sets_match_reference(oracle_mod, name):
    from pagerank_amd import apply_batch
    z, x, b = _load(name)
    y = apply_batch(x, b)
    assert np.array_equal(oracle_mod.mark(oracle_mod.DYNAMIC_FRONTIER, y, b), z["affected_df"])
    assert np.array_equal(oracle_mod.mark(oracle_mod.DYNAMIC_TRAVERSAL, y, b), z["affected_dt"])


@pytest.mark.parametrize("name", FIXTURES)
def test_frontier_methods_reach_reference_solution(oracle_mod, name):
    """The CPU DF / DF-P mark neighbours during the sweep and the GPU path after it, so their iterates differ; both
    must land within the tolerance-level distance of each other and of the exact fixed point."""
    from pagerank_amd import apply_batch
    z, x, b = _load(name)
    y = apply_batch(x, b)
    exact, _ = oracle_mod.power_iteration(y)
    r, _, _ = oracle_mod.pagerank_cuda(oracle_mod.DYNAMIC_FRONTIER, y, b, z["q"])
    assert np.abs(r - exact).max() <= 1e-9
    assert np.abs(z["DF_ranks"] - exact).max() <= 1e-9
    r, _, _ = oracle_mod.pagerank_cuda(oracle_mod.PRUNE_DYNAMIC_FRONTIER, y, b, z["q"])
    assert np.abs(r - exact).max() <= 1e-5 * exact.max()
    assert np.abs(z["DFP_ranks"] - exact).max() <= 1e-5 * exact.max()


@pytest.mark.parametrize("make", [lambda: cycle_graph(257), lambda: complete_graph(33), lambda: torus_graph(9, 7),
                                  lambda: star_graph(300)], ids=["cycle", "complete", "torus", "star"])
def test_closed_forms_and_invariants(oracle_mod, make):
    y = make()
    r, it, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y)
    assert abs(r.sum() - 1) < 1e-12 and r.min() >= 0.15 / y.N * (1 - 1e-12)
    if y.N != 300:  # vertex-transitive: uniform up to the rounding of the N-term sum, one iteration
        assert it == 1 and np.abs(r * y.N - 1.0).max() <= 1e-14
    else:  # star: hub rank from the 2x2 reduced system of the symmetric star
        n = y.N
        # leaves: l = (1-a)/n + a*(l/2 + h/n) with hub out-degree n; hub: h = (1-a)/n + a*(h/n + (n-1) l/2)
        a = 0.85
        A = np.array([[1 - a / 2, -a / n], [-a * (n - 1) / 2, 1 - a / n]])
        l, h = np.linalg.solve(A, np.full(2, (1 - a) / n))
        assert abs(r[0] - h) <= 1e-9 and np.abs(r[1:] - l).max() <= 1e-9


def test_dense_solve(oracle_mod):
    from pagerank_amd import HostGraph
    y = HostGraph.rmat(9, 8, seed=5)
    exact = np.linalg.solve(np.eye(y.N) - 0.85 * y.dense_transition(), np.full(y.N, 0.15 / y.N))
    p, _ = oracle_mod.power_iteration(y)
    assert np.abs(p - exact).max() <= 1e-15
    r, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y, tolerance=1e-15)
    assert np.abs(r - exact).max() <= 1e-14


def test_dead_ends(oracle_mod):
    """A graph with dead ends (out-degree 0, no self-loops): the restatement solves (I - a A) r = (1-a)/N 1 with the
    dead ends' columns of A zero (numpy.linalg.solve), and, where oracle/_ref is built, the OpenMP port equals the
    compiled pagerankStaticOmp bit for bit."""
    from pagerank_amd import HostGraph
    x = HostGraph.rmat(9, 4, seed=21, self_loops=False)
    src = np.repeat(np.arange(x.N, dtype=np.uint32), np.diff(x.out_off.astype(np.int64)))
    keep = (src % 3) != 0
    y = HostGraph.from_edges(x.N, src[keep], x.out_edg[keep], self_loops=False)
    assert int((y.outdeg == 0).sum()) >= x.N // 3
    A = np.zeros((y.N, y.N))
    for u in range(y.N):
        for v in y.out_edg[y.out_off[u]:y.out_off[u + 1]]:
            A[v, u] += 1.0 / y.outdeg[u]
    exact = np.linalg.solve(np.eye(y.N) - 0.85 * A, np.full(y.N, 0.15 / y.N))
    r, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y, tolerance=1e-15)
    assert np.abs(r - exact).max() <= 1e-14 and exact.sum() < 1.0
    if oracle_mod.ref_available():
        rr, itr, _ = oracle_mod.ref_pagerank(0, y)
        ro, ito, _ = oracle_mod.static_omp(y)
        assert itr == ito and np.array_equal(rr, ro)


@pytest.mark.parametrize("n", [1, 5, 64, 3000])
def test_partition_brute_force(oracle_mod, n):
    rng = np.random.default_rng(n)
    deg = rng.integers(0, 130, size=n)
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    order, npart = oracle_mod.partition(off, 64)
    want = [v for v in range(n) if deg[v] <= 64] + [v for v in range(n) if deg[v] > 64]
    assert npart == int((deg <= 64).sum()) and order.tolist() == want


def test_traversal_brute_force(oracle_mod):
    from pagerank_amd import Batch, HostGraph, apply_batch
    x = HostGraph.rmat(9, 2, seed=13)
    b = Batch.random(x, 6, 0.5, seed=2)
    y = apply_batch(x, b)
    got = oracle_mod.mark(oracle_mod.DYNAMIC_TRAVERSAL, y, b)
    seen = np.zeros(y.N, bool)
    stack = [int(v) for v in b.del_v] + [int(w) for u in np.concatenate([b.del_u, b.ins_u])
                                         for w in y.out_edg[y.out_off[u]:y.out_off[u + 1]]]
    while stack:
        v = stack.pop()
        if seen[v]:
            continue
        seen[v] = True
        stack.extend(int(w) for w in y.out_edg[y.out_off[v]:y.out_off[v + 1]])
    assert np.array_equal(got.astype(bool), seen)


def test_live_reference_agrees(oracle_mod):
    """Where oracle/_ref is built (this container), also check a fresh graph against it directly."""
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built here (the reference sources are only present in the build container)")
    from pagerank_amd import Batch, HostGraph, apply_batch
    x = HostGraph.rmat(11, 16, seed=99)
    b = Batch.random(x, 300, 0.8, seed=4)
    y = apply_batch(x, b)
    rr, itr, _ = oracle_mod.ref_pagerank(0, x, b)
    r, it, _ = oracle_mod.static_omp(y)
    assert it == itr and np.array_equal(r, rr)
    assert np.array_equal(oracle_mod.mark(2, y, b), oracle_mod.ref_affected(1, x, b))
    q, _, _ = oracle_mod.ref_pagerank(0, x)
    for m in (1, 4):
        rr, itr, _ = oracle_mod.ref_pagerank(m, x, b, q, asynchronous=True, threads=1)
        r, it, _ = oracle_mod.pagerank_async(m, y, b, q, threads=1)
        assert it == itr and np.array_equal(r, rr), m


def test_pins_reject_planted_mistakes(oracle_mod):
    """Each planted mistake in the oracle's inputs must fail at least one pin."""
    from pagerank_amd import HostGraph
    y = HostGraph.rmat(8, 8, seed=21)
    exact = np.linalg.solve(np.eye(y.N) - 0.85 * y.dense_transition(), np.full(y.N, 0.15 / y.N))

    def pin(g):
        r, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, g, tolerance=1e-14)
        return np.abs(r - exact).max() <= 1e-12

    assert pin(y)
    # transposed operand: pull over out-neighbours instead of in-neighbours
    t = HostGraph(arrays=dict(out_off=y.out_off, out_edg=y.out_edg, in_off=y.out_off, in_edg=y.out_edg,
                              in_eid=y.in_eid, outdeg=y.outdeg))
    assert not pin(t)
    # off-by-one source index
    s = HostGraph(arrays=dict(out_off=y.out_off, out_edg=y.out_edg, in_off=y.in_off,
                              in_edg=(y.in_edg.astype(np.int64) + 1) % y.N, in_eid=y.in_eid, outdeg=y.outdeg))
    assert not pin(s)
    # wrong per-source divisor (in-degree instead of out-degree)
    d = HostGraph(arrays=dict(out_off=y.out_off, out_edg=y.out_edg, in_off=y.in_off, in_edg=y.in_edg,
                              in_eid=y.in_eid, outdeg=np.diff(y.in_off.astype(np.int64))))
    assert not pin(d)
    # dropped teleport term / wrong damping
    r, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y, damping=0.86, tolerance=1e-14)
    assert np.abs(r - exact).max() > 1e-9


def _dense_weighted(y, w, s, c, a=0.85):
    """Dense FP64 forward solve and torch autograd gradient of L = sum c r^2 (independent of the adjoint formula)."""
    torch = pytest.importorskip("torch")
    src = torch.tensor(y.out_sources().astype(np.int64))
    dst = torch.tensor(y.out_edg.astype(np.int64))
    wt = torch.tensor(w.astype(np.float64), requires_grad=True)
    st = torch.tensor(s, dtype=torch.float64, requires_grad=True)
    Wt = torch.zeros(y.N, dtype=torch.float64).index_add(0, src, wt)
    P = torch.zeros(y.N, y.N, dtype=torch.float64).index_put((src, dst), wt / Wt[src], accumulate=True)
    r = torch.linalg.solve(torch.eye(y.N, dtype=torch.float64) - a * P.T, (1 - a) * st)
    (torch.tensor(c) * r * r).sum().backward()
    return r.detach().numpy(), wt.grad.numpy(), st.grad.numpy()


def test_weighted_oracle_matches_dense_autograd(oracle_mod):
    """The oracle's weighted forward, adjoint and gradient assembly against a dense FP64 solve differentiated by
    torch autograd on a 2,048-vertex RMAT graph (run to stagnation)."""
    from pagerank_amd import HostGraph
    y = HostGraph.rmat(11, 16, seed=5)
    rng = np.random.default_rng(0)
    w = rng.uniform(0.5, 1.5, y.M).astype(np.float32)
    c = rng.standard_normal(y.N)
    s = np.full(y.N, 1.0 / y.N)
    r, _, W = oracle_mod.weighted_forward(y, w, s, tolerance=1e-15, max_iterations=2000)
    g = 2 * c * r
    lam, _ = oracle_mod.weighted_adjoint(y, w, W, g, tolerance=1e-15, max_iterations=2000)
    gw, gs = oracle_mod.weighted_grad(y, W, r, lam, g)
    rd, gwd, gsd = _dense_weighted(y, w, s, c)
    assert np.abs(r - rd).max() <= 1e-13
    assert np.abs(gw - gwd).max() <= 1e-15
    assert np.abs(gs - gsd).max() <= 1e-12
    # a planted mistake (dropping the normalisation term of dP/dw) must fail the pin
    wrong = gw + np.repeat(((lam - g) / 0.85) * 0.85 * r / W, np.diff(y.out_off.astype(np.int64)))
    assert np.abs(wrong - gwd).max() > 1e-9


def test_ppr_oracle_matches_dense_solve(oracle_mod):
    """Personalized PageRank r = (1-alpha) s + alpha A r, solved densely (numpy.linalg.solve), for a seed-set teleport
    and for a point mass."""
    from pagerank_amd import HostGraph
    y = HostGraph.rmat(8, 8, seed=13)
    A = y.dense_transition()
    rng = np.random.default_rng(5)
    for seeds in (rng.choice(y.N, 17, replace=False), np.array([3])):
        s = np.zeros(y.N)
        s[seeds] = 1.0 / len(seeds)
        exact = np.linalg.solve(np.eye(y.N) - 0.85 * A, 0.15 * s)
        r, _ = oracle_mod.ppr(y, s, tolerance=1e-14, max_iterations=2000)
        assert np.abs(r - exact).max() <= 1e-12
        assert abs(r.sum() - 1.0) <= 1e-12


def test_dfp_update_matches_compiled_reference(oracle_mod):
    """One DF-P iteration from q: on the initially affected vertices, the oracle's update equals the compiled
    reference's pagerankPruneUpdateRank (pagerankPrune.hxx:39) up to FMA rounding. The CPU path marks neighbours
    during its sweep, so the comparison is restricted to the vertices both mark before the first update."""
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built here (the reference sources are only present in the build container)")
    from pagerank_amd import Batch, HostGraph, apply_batch
    x = HostGraph.rmat(10, 8, seed=17)
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    b = Batch.random(x, 40, 0.8, seed=2)
    y = apply_batch(x, b)
    rr, itr, _ = oracle_mod.ref_pagerank(oracle_mod.PRUNE_DYNAMIC_FRONTIER, x, b, q, max_iterations=1)
    r, it, vaff = oracle_mod.pagerank_cuda(oracle_mod.PRUNE_DYNAMIC_FRONTIER, y, b, q, max_iterations=1)
    init = vaff.astype(bool)
    assert it == itr == 1 and init.sum() > 10
    moved = np.abs(r - q) > 0
    assert moved[init].mean() > 0.5  # the update did change the affected ranks
    assert np.abs(r[init] - rr[init]).max() <= 1e-15 * np.abs(rr[init]).max()
    # a plain Jacobi update (the self-loop term not solved in closed form) would fail the same comparison
    j, _, _ = oracle_mod.pagerank_cuda(oracle_mod.DYNAMIC_FRONTIER, y, b, q, max_iterations=1)
    assert np.abs(j[init] - rr[init]).max() > 1e-6 * np.abs(rr[init]).max()
