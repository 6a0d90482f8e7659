"""Weighted PageRank, its adjoint and the torch.autograd harness on the GPU.

Bars: forward ranks, adjoint lambda and both gradients within L-inf 1e-9 of the oracle at the same tolerance
(1e-10); at a tight tolerance the GPU gradient matches a dense FP64 solve differentiated by torch autograd on a
2,048-vertex RMAT graph (1e-15 on dL/dw, 1e-12 on dL/ds); the autograd harness's dL/dw matches central finite
differences (h = 1e-6 on FP32 weights) on 32 random edges with forward solves run to stagnation.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(scale=11, seed=5):
    from pagerank_amd import HostGraph
    y = HostGraph.rmat(scale, 16, seed=seed)
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 1.5, y.M).astype(np.float32)
    c = rng.standard_normal(y.N)
    s = np.full(y.N, 1.0 / y.N)
    return y, w, c, s


def test_uniform_weights_equal_static(ctx):
    from pagerank_amd import HostGraph, pagerankStaticHip, weighted_pagerank
    y = HostGraph.rmat(12, 16, seed=3)
    yg = ctx.upload(y)
    r = weighted_pagerank(yg, np.ones(y.M, np.float32), np.full(y.N, 1.0 / y.N))
    st = pagerankStaticHip(yg)
    assert abs(r.iterations - st.iterations) <= 1
    assert np.abs(r.ranks - st.ranks).max() <= 1e-9


def _star_case(n=12000, seed=6):
    from conftest import star_graph
    y = star_graph(n)
    rng = np.random.default_rng(seed)
    return y, rng.uniform(0.5, 1.5, y.M).astype(np.float32), rng.standard_normal(y.N), rng.dirichlet(np.ones(y.N))


@pytest.mark.parametrize("blocking", [0, 1024, 4096], ids=["rows", "blocked", "blocked4096"])
@pytest.mark.parametrize("case", ["rmat11", "rmat14", "star12000"])
def test_forward_and_adjoint_match_oracle(ctx, oracle_mod, case, blocking):
    """rmat11/rmat14: rows of every degree class; star12000: a hub whose in-row (forward) and out-row (adjoint)
    hold 12000 edges, summed in chunks of 4096 by separate workgroups and combined by a second launch. blocked: the
    solves run through the propagation-blocked plan (B = 1024), the edge weight applied in expand."""
    from pagerank_amd import Tuning
    y, w, c, s = _star_case() if case.startswith("star") else _case(int(case[4:]))
    print(f"[adjoint] case={case} blocking={blocking} N={y.N} M={y.M}", flush=True)
    old = ctx.get_tuning()
    ctx.set_tuning(Tuning(**{**old.__dict__, "blocking": blocking}))
    try:
        _forward_and_adjoint(ctx, oracle_mod, y, w, c, s)
    finally:
        ctx.set_tuning(old)


def _forward_and_adjoint(ctx, oracle_mod, y, w, c, s):
    from pagerank_amd import weighted_adjoint, weighted_pagerank
    yg = ctx.upload(y)
    fw = weighted_pagerank(yg, w, s)
    r, it, W = oracle_mod.weighted_forward(y, w, s)
    assert abs(fw.iterations - it) <= 1
    assert np.abs(fw.ranks - r).max() <= 1e-9
    g = 2 * c * fw.ranks
    lam, gw, gs, ait = weighted_adjoint(yg, w, fw.ranks, g)
    lam_o, ait_o = oracle_mod.weighted_adjoint(y, w, W, g)
    gw_o, gs_o = oracle_mod.weighted_grad(y, W, fw.ranks, lam_o, g)
    assert abs(ait - ait_o) <= 1
    assert np.abs(lam - lam_o).max() <= 1e-9
    assert np.abs(gw - gw_o).max() <= 1e-9
    assert np.abs(gs - gs_o).max() <= 1e-9


def test_gradient_matches_dense_autograd(ctx):
    torch = pytest.importorskip("torch")
    from pagerank_amd import PagerankOptions, weighted_adjoint, weighted_pagerank
    y, w, c, s = _case(11)
    yg = ctx.upload(y)
    o = PagerankOptions(tolerance=1e-15, maxIterations=2000)
    fw = weighted_pagerank(yg, w, s, o)
    lam, gw, gs, _ = weighted_adjoint(yg, w, fw.ranks, 2 * c * fw.ranks, o)
    src = torch.tensor(y.out_sources().astype(np.int64))
    dst = torch.tensor(y.out_edg.astype(np.int64))
    wt = torch.tensor(w.astype(np.float64), requires_grad=True)
    st = torch.tensor(s, requires_grad=True)
    Wt = torch.zeros(y.N, dtype=torch.float64).index_add(0, src, wt)
    P = torch.zeros(y.N, y.N, dtype=torch.float64).index_put((src, dst), wt / Wt[src], accumulate=True)
    r = torch.linalg.solve(torch.eye(y.N, dtype=torch.float64) - 0.85 * P.T, 0.15 * st)
    (torch.tensor(c) * r * r).sum().backward()
    assert np.abs(fw.ranks - r.detach().numpy()).max() <= 1e-13
    assert np.abs(gw - wt.grad.numpy()).max() <= 1e-15
    assert np.abs(gs - st.grad.numpy()).max() <= 1e-12


def test_autograd_harness_finite_differences(ctx):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch has no GPU here")
    from pagerank_amd import PagerankOptions
    from pagerank_amd.autograd import pagerank
    y, w, c, s = _case(11, seed=8)
    yg = ctx.upload(y)
    dev = torch.device("cuda:0")
    cw = torch.tensor(c, device=dev)
    wt = torch.tensor(w, device=dev, requires_grad=True)
    st = torch.tensor(s, device=dev, requires_grad=True)
    o = PagerankOptions(tolerance=1e-10)
    loss = (cw * pagerank(wt, st, yg, o) ** 2).sum()
    loss.backward()
    grad = wt.grad.double().cpu().numpy()
    assert wt.grad.dtype == torch.float32 and st.grad.dtype == torch.float64
    stag = PagerankOptions(tolerance=0.0, maxIterations=400)  # to stagnation: 0.85^400 ~ 1e-28

    def L(wv):
        with torch.no_grad():
            return float((cw * pagerank(torch.tensor(wv, device=dev), st.detach(), yg, stag) ** 2).sum())

    rng = np.random.default_rng(3)
    for e in rng.choice(y.M, size=32, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[e] = np.float32(w[e] + np.float32(1e-6))
        wm[e] = np.float32(w[e] - np.float32(1e-6))
        fd = (L(wp) - L(wm)) / (float(wp[e]) - float(wm[e]))
        assert abs(fd - grad[e]) <= 1e-10 + 1e-4 * abs(fd), (e, fd, grad[e])


def test_fp64_inputs_get_fp64_gradients(ctx):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch has no GPU here")
    from pagerank_amd import weighted_adjoint, weighted_pagerank
    from pagerank_amd.autograd import pagerank
    y, w, c, s = _case(10, seed=2)
    yg = ctx.upload(y)
    dev = torch.device("cuda:0")
    wt = torch.tensor(w.astype(np.float64), device=dev, requires_grad=True)
    st = torch.tensor(s, device=dev, requires_grad=True)
    r = pagerank(wt, st, yg)
    (torch.tensor(c, device=dev) * r * r).sum().backward()
    fw = weighted_pagerank(yg, w, s)
    _, gw, gs, _ = weighted_adjoint(yg, w, fw.ranks, 2 * c * fw.ranks)
    assert wt.grad.dtype == torch.float64
    assert np.array_equal(r.detach().cpu().numpy(), fw.ranks)
    assert np.abs(wt.grad.cpu().numpy() - gw).max() <= 1e-15
    assert np.abs(st.grad.cpu().numpy() - gs).max() <= 1e-15
