"""The autotuner (prh_autotune): picks one of its candidate launch plans, caches the choice per device and graph
in a JSON file, and the tuned context still gives the oracle's ranks."""
from __future__ import annotations

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_choice_cache_and_parity(ctx, oracle_mod, tmp_path):
    from pagerank_amd import HostGraph, pagerankStaticHip
    cache = tmp_path / "tuning.json"
    old = ctx.get_tuning()
    try:
        g1 = HostGraph.rmat(12, 16, seed=5)
        yg1 = ctx.upload(g1, with_out=False)
        print("[autotune] first graph", flush=True)
        t1, ms1, hit1 = ctx.autotune(yg1, str(cache), iterations=4)
        assert not hit1 and ms1 > 0
        assert t1.source_slices in (0, 1, 8) and t1.switch_degree in (64, 128, 256)
        assert t1.huge_degree in (4096, 16384) and t1.block_edges in (2048, 4096)
        assert ctx.get_tuning() == t1
        entries = json.loads(cache.read_text())
        assert len(entries) == 1
        (key, rec), = entries.items()
        assert key.startswith("gfx950") and f"-n{g1.N}-m{g1.M}-" in key
        assert rec["source_slices"] == t1.source_slices and rec["ms_per_iteration"] > 0
        res = pagerankStaticHip(yg1)
        ref, it, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, g1)
        assert abs(res.iterations - it) <= 1 and np.abs(res.ranks - ref).max() <= 1e-9
        # a hit: same choice, no timing
        ctx.set_tuning(old)
        t1b, ms1b, hit1b = ctx.autotune(yg1, str(cache), iterations=4)
        assert hit1b and t1b == t1 and ms1b == pytest.approx(rec["ms_per_iteration"])
        # another graph adds an entry
        g2 = HostGraph.grid(40, 30, 0.62, seed=2)
        yg2 = ctx.upload(g2, with_out=False)
        print("[autotune] second graph", flush=True)
        _, _, hit2 = ctx.autotune(yg2, str(cache), iterations=4)
        assert not hit2 and len(json.loads(cache.read_text())) == 2
        res2 = pagerankStaticHip(yg2)
        ref2, it2, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, g2)
        assert np.abs(res2.ranks - ref2).max() <= 1e-9
    finally:
        ctx.set_tuning(old)


def test_autotune_large_graph_considers_blocking(ctx, oracle_mod, tmp_path):
    """From 2^20 edges on, the propagation-blocked plan (B = 4096, 8192, 16384) is among the candidates; whichever
    plan wins, the cache records its block width and the tuned solve gives the oracle's ranks."""
    from pagerank_amd import HostGraph, pagerankStaticHip
    cache = tmp_path / "tuning.json"
    old = ctx.get_tuning()
    try:
        g = HostGraph.rmat(16, 20, seed=13)
        assert g.M >= 1 << 20
        yg = ctx.upload(g, with_out=False)
        print(f"[autotune-large] N={g.N} M={g.M}", flush=True)
        t, ms, hit = ctx.autotune(yg, str(cache), iterations=4)
        assert not hit and ms > 0 and t.blocking in (0, 4096, 8192, 16384)
        (rec,) = json.loads(cache.read_text()).values()
        assert rec["blocking"] == t.blocking
        res = pagerankStaticHip(yg)
        ref, it, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, g)
        assert abs(res.iterations - it) <= 1 and np.abs(res.ranks - ref).max() <= 1e-9
    finally:
        ctx.set_tuning(old)
