"""The C++ drop-in header on the GPU: callers of the reference's five entry points (pagerankCuda.hxx:864-988),
compiled against host/pagerankHip.hxx, give the oracle's results.

* tests/cpp/_build/dropin_selftest -- the package's CSR graph class (always built by the package Makefile);
* oracle/_ref/dropin_main -- the reference's own DiGraph, batch and CPU headers with pagerankHip.hxx in place of
  pagerankCuda.hxx (built where the reference sources exist; skipped otherwise).
Bars: L-inf <= 1e-9 against the oracle's restatement of the same solver (DF-P: against its own restatement, plus
its approximation bound), and the reference experiment's L1 error metric (main.cu:55) below 1e-8 for the exact
methods.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = Path(__file__).resolve().parent.parent
SELFTEST = REPO / "tests" / "cpp" / "_build" / "dropin_selftest"
REF_DROPIN = REPO / "oracle" / "_ref" / "dropin_main"


def _case(tmp_path):
    from pagerank_amd import Batch, HostGraph, apply_batch
    x = HostGraph.rmat(11, 16, seed=77)
    b = Batch.random(x, 150, 0.8, seed=6)
    y = apply_batch(x, b)
    x.out_off.tofile(tmp_path / "x_off.bin")
    x.out_edg.tofile(tmp_path / "x_edg.bin")
    y.out_off.tofile(tmp_path / "y_off.bin")
    y.out_edg.tofile(tmp_path / "y_edg.bin")
    np.concatenate([[b.deletions, b.insertions], b.del_u, b.del_v, b.ins_u, b.ins_v]).astype(np.uint32).tofile(
        tmp_path / "batch.bin")
    return x, b, y


def test_selftest_program(tmp_path, oracle_mod):
    if not SELFTEST.exists():
        pytest.fail(f"{SELFTEST} not built (make -C pagerank-cuda-dynamic_amd)")
    x, b, y = _case(tmp_path)
    q, _, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, x)
    q.tofile(tmp_path / "q.bin")
    p = subprocess.run([str(SELFTEST)] + [str(tmp_path / f) for f in ("x_off.bin", "x_edg.bin", "y_off.bin",
                                                                        "y_edg.bin", "batch.bin", "q.bin")]
                       + [str(tmp_path / "out")], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    lines = {ln.split()[0]: ln for ln in p.stdout.splitlines()}
    for name, method in (("static", 0), ("naive", 1), ("traversal", 2), ("frontier", 3), ("prune", 4)):
        got = np.fromfile(tmp_path / f"out.{name}.bin")
        ref, it, _ = oracle_mod.pagerank_cuda(method, y, b, q)
        iters = int(lines[name].split("iterations=")[1].split()[0])
        assert abs(iters - it) <= 1, (name, iters, it)
        assert np.abs(got - ref).max() <= 1e-9, name
    assert "empty ranks=0 iterations=0" in p.stdout
    assert "unknown-vertex: rejected" in p.stdout


def test_reference_caller_compiled_against_header(tmp_path, oracle_mod):
    if not REF_DROPIN.exists():
        pytest.skip("oracle/_ref/dropin_main is built only where the reference sources are present")
    x, b, y = _case(tmp_path)
    p = subprocess.run([str(REF_DROPIN), str(tmp_path / "x_off.bin"), str(tmp_path / "x_edg.bin"),
                        str(tmp_path / "batch.bin"), str(tmp_path / "out")], capture_output=True, text=True,
                       timeout=300)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    q = np.fromfile(tmp_path / "out.q.bin")  # the reference's r0 = pagerankStaticOmp(xt, tolerance 1e-100)
    methods = {"static": 0, "static_partition": 0, "naive": 1, "traversal": 2, "frontier": 3, "prune": 4}
    seen = set()
    for line in p.stdout.splitlines():
        name = line.split()[0]
        seen.add(name)
        err = float(line.split("l1err=")[1].split()[0])
        got = np.fromfile(tmp_path / f"out.{name}.bin")
        assert len(got) == y.N
        ref, _, _ = oracle_mod.pagerank_cuda(methods[name], y, b, q)
        assert np.abs(got - ref).max() <= 1e-9, name
        # L1 error against the reference's converged CPU ranks (main.cu:55): tolerance-level for the exact methods
        assert err < (1e-4 if name == "prune" else 1e-6), line
    assert seen == set(methods)


def _run_static(tmp_path, y, repeats: int, deterministic: bool):
    import os
    y.out_off.tofile(tmp_path / "y_off.bin")
    y.out_edg.tofile(tmp_path / "y_edg.bin")
    env = {k: v for k, v in os.environ.items() if k not in ("PRH_DETERMINISTIC", "PRH_TUNING_CACHE")}
    if deterministic:
        env["PRH_DETERMINISTIC"] = "1"
    p = subprocess.run([str(SELFTEST), "--static", str(tmp_path / "y_off.bin"), str(tmp_path / "y_edg.bin"),
                        str(repeats), str(tmp_path / "s")], capture_output=True, text=True, timeout=300, env=env)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    rows = []
    for line in p.stdout.splitlines():
        f = dict(kv.split("=") for kv in line.split()[1:])
        rows.append({k: float(v) for k, v in f.items()})
    return rows, [np.fromfile(tmp_path / f"s.{i}.bin") for i in range(repeats)]


def test_dropin_large_graph_takes_autotuned_plan(tmp_path, oracle_mod):
    """A graph of 2^20 edges or more: the drop-in runs the autotuner's plan (the propagation-blocked update on
    RMAT-20), reports its one-time build in the first solve's initialization time, and matches the oracle."""
    if not SELFTEST.exists():
        pytest.fail(f"{SELFTEST} not built (make -C pagerank-cuda-dynamic_amd)")
    from pagerank_amd import HostGraph
    y = HostGraph.rmat(20, 16, seed=5)
    rows, ranks = _run_static(tmp_path, y, 2, deterministic=False)
    assert all(r["blocking"] > 0 for r in rows), rows
    ref, it, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y)
    for r, got in zip(rows, ranks):
        assert abs(int(r["iterations"]) - it) <= 1
        assert np.abs(got - ref).max() <= 1e-9
        assert r["init"] >= r["time"] - r["comp"] - 0.05  # time = init + comp (+ an empty marking phase)


def test_dropin_deterministic_keeps_row_plan(tmp_path, oracle_mod):
    """PRH_DETERMINISTIC=1: the row plan, whose fixed summation order repeats bit for bit."""
    if not SELFTEST.exists():
        pytest.fail(f"{SELFTEST} not built (make -C pagerank-cuda-dynamic_amd)")
    from pagerank_amd import HostGraph
    y = HostGraph.rmat(20, 16, seed=5)
    rows, ranks = _run_static(tmp_path, y, 2, deterministic=True)
    assert all(r["blocking"] == 0 for r in rows), rows
    assert np.array_equal(ranks[0], ranks[1])
    ref, it, _ = oracle_mod.pagerank_cuda(oracle_mod.STATIC, y)
    assert abs(int(rows[0]["iterations"]) - it) <= 1
    assert np.abs(ranks[0] - ref).max() <= 1e-9
