"""The experiment runner end to end on the GPU: temporal mode on a small SNAP temporal file (main.cu:157-174's
protocol) and dynamic mode on a small RMAT graph. Checks the graph sizes each batch reaches (recomputed with numpy
from the file) and the reference experiment's L1 error metric (main.cu:55) for every method: at most 1e-9 per vertex
for the methods that iterate to tau = 1e-10 in L-inf (with no dead ends the iteration contracts by a in L1, so the
last step's L1 change, at most N * tau, bounds the L1 distance to the fixed point by N * tau * a / (1 - a), about
5.7e-10 per vertex), and 1e-8 per vertex for DF-P, whose pruned vertices stop early."""
from __future__ import annotations

import csv

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXACT = {"pagerankStaticHip", "pagerankNaiveDynamicHip", "pagerankDynamicTraversalHip",
         "pagerankDynamicFrontierHip"}


def _bound(r) -> float:
    return int(r["vertices"]) * (1e-9 if r["technique"] in EXACT else 1e-8)


def _rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def test_temporal_mode(tmp_path, capsys):
    from pagerank_amd import experiment
    rng = np.random.default_rng(5)
    rows, size = 600, 20000
    u = (rng.zipf(1.6, size) % rows) + 1
    v = rng.integers(1, rows + 1, size)
    p = tmp_path / "t.txt"
    p.write_text("".join(f"{a} {b} {1000 + i}\n" for i, (a, b) in enumerate(zip(u.tolist(), v.tolist()))))
    out = tmp_path / "r.csv"
    assert experiment.main(["temporal", "--file", str(p), "--rows", str(rows), "--size", str(size),
                            "--fractions", "1e-3", "--batches", "3", "--repeat", "1", "--no-autotune",
                            "--csv", str(out)]) == 0
    got = _rows(out)
    assert len(got) == 3 * 5
    base = int(0.9 * size)
    for bi in range(3):
        k = base + (bi + 1) * 20
        keys = np.unique((u[:k] - 1).astype(np.int64) * rows + (v[:k] - 1))
        keys = np.union1d(keys, np.arange(rows, dtype=np.int64) * (rows + 1))
        for r in (r for r in got if int(r["batch_index"]) == bi):
            assert int(r["vertices"]) == rows and int(r["edges"]) == len(keys)
            assert float(r["l1_error"]) < _bound(r), r
    assert "pagerankPruneDynamicFrontierHip" in capsys.readouterr().out


@pytest.mark.parametrize("path", ["update", "upload"])
def test_dynamic_mode(tmp_path, path):
    """Both ways of making each updated graph: the device update (default) and the host build + upload. Every
    method's time includes the batch's preparation."""
    from pagerank_amd import experiment
    out = tmp_path / "r.csv"
    extra = ["--upload"] if path == "upload" else []
    assert experiment.main(["dynamic", "--scale", "12", "--fractions", "1e-4,1e-3", "--batches", "2",
                            "--repeat", "1", "--csv", str(out)] + extra) == 0
    got = _rows(out)
    assert len(got) == 2 * 2 * 5
    for r in got:
        assert float(r["l1_error"]) < _bound(r), r
        prep = float(r["extra"].split("preparation_ms=")[1].split(";")[0])
        assert prep > 0 and float(r["init_ms"]) >= prep and r["extra"].endswith(path)


def test_adjoint_mode(tmp_path):
    """Adjoint mode on a small RMAT graph: the host-array pair and the torch.autograd pair on device tensors both
    run and report their times."""
    from pagerank_amd import experiment
    out = tmp_path / "a.csv"
    assert experiment.main(["adjoint", "--scale", "12", "--repeat", "1", "--csv", str(out)]) == 0
    rows = {r["technique"]: r for r in _rows(out)}
    assert set(rows) == {"weighted_forward", "weighted_adjoint_wall", "autograd_forward_wall",
                         "autograd_backward_wall", "static"}
    assert all(float(r["time_ms"]) > 0 for r in rows.values())
    assert int(rows["autograd_backward_wall"]["iterations"]) == int(rows["weighted_adjoint_wall"]["iterations"])
