"""The dynamic configuration at its stated size: directed RMAT scale 26 with edge factor 30 (67.1M vertices, 2.04B
edges with self-loops, close to the 2^32 offset limit) and one update batch of 1e-5 |E| (80% insertions, 20%
deletions), applied to the device graph in place (prh_graph_update).

Checks, against the oracle on the host-applied graph (the reference checks every solver against a converged CPU
solve, main.cu:55, :88):
  * the updated graph's out-CSR equals the host-applied graph's (apply_batch) bit for bit;
  * static PageRank (autotuned plan, patched in place) against the OpenMP port of pagerankStaticOmp: L-inf <= 1e-9,
    iterations within 1, ranks summing to 1;
  * the DT affected set bit-exact with the oracle's traversal closure;
  * DF-P from the previous graph's ranks within 1e-5 of the largest rank of the port's solution (DF-P prunes).
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import emit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3(request):
    from pagerank_amd import Batch, Context, HostGraph, apply_batch, solve
    say = lambda m: emit(request.config, f"C3: {m}")  # noqa: E731
    t0 = time.time()
    x = HostGraph.rmat(26, 30, 0.57, 0.19, 0.19, seed=42)
    say(f"RMAT-26 x30: N={x.N} M={x.M} built in {time.time() - t0:.0f} s")
    b = Batch.random(x, int(1e-5 * x.M), 0.8, seed=7)
    y = apply_batch(x, b)
    say(f"batch -{b.deletions}/+{b.insertions}, host-applied y: M={y.M} ({time.time() - t0:.0f} s)")
    ctx = Context(0)
    xg = ctx.upload(x)
    t, ms, _ = ctx.autotune(xg, None, iterations=3)
    q = solve(xg, 0).ranks
    upd = xg.update(b)
    say(f"upload, autotune (blocking {t.blocking}, {ms:.2f} ms/iteration), static of x, update {upd:.1f} ms "
        f"({time.time() - t0:.0f} s)")
    yield dict(x=x, y=y, b=b, ctx=ctx, g=xg, q=q, tuning=t)
    xg.close()
    ctx.close()


def test_c3_updated_graph_equals_host_applied(c3):
    off, idx = c3["g"].csr("out")
    assert np.array_equal(off, c3["y"].out_off) and np.array_equal(idx, c3["y"].out_edg)


def test_c3_static_against_port(c3, oracle_mod, say):
    from pagerank_amd import solve
    res = solve(c3["g"], 0)
    say(f"GPU static on y: {res.iterations} iterations, {res.computationTime:.1f} ms; oracle port ...")
    ref, it, cpu_ms = oracle_mod.static_omp(c3["y"])
    err = float(np.abs(res.ranks - ref).max())
    say(f"port: {it} iterations in {cpu_ms / 1e3:.1f} s, L-inf {err:.3e}")
    assert abs(res.iterations - it) <= 1
    assert err <= 1e-9
    assert abs(res.ranks.sum() - 1.0) <= 1e-9
    c3["port"] = ref


def test_c3_traversal_set_bit_exact(c3, oracle_mod, say):
    got = c3["ctx"].mark_affected(c3["g"], 2, c3["b"])
    want = oracle_mod.mark(2, c3["y"], c3["b"])
    say(f"DT affected: {int(got.sum())} of {len(got)} vertices (oracle {int(want.sum())})")
    assert np.array_equal(got, want)


def test_c3_dfp_within_bound(c3, oracle_mod, say):
    from pagerank_amd import solve
    res = solve(c3["g"], 4, batch=c3["b"], q=c3["q"])
    ref = c3.get("port")
    if ref is None:
        ref, _, _ = oracle_mod.static_omp(c3["y"])
    err = float(np.abs(res.ranks - ref).max())
    say(f"DF-P: {res.iterations} iterations, {res.time:.1f} ms; L-inf vs the port {err:.3e}")
    assert err <= 1e-5 * ref.max()
