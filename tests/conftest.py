"""Shared test setup: import paths, the `gpu` marker, and small graph fixtures."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "pagerank-cuda-dynamic_amd"))
sys.path.insert(0, str(REPO / "oracle"))

# torch bundles its own HIP runtime under the same soname as the system one the engine links; whichever loads first
# serves the process. Load torch's first so the autograd-harness tests and the engine share one runtime.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP engine library")


def emit(config, msg: str) -> None:
    """A progress line that reaches the terminal while a long test or fixture runs (output capture is suspended for
    the line), so that a run of several minutes shows it is alive."""
    capman = config.pluginmanager.getplugin("capturemanager")
    if capman is None:
        print(msg, flush=True)
        return
    with capman.global_and_fixture_disabled():
        print(f"\n[progress] {msg}", flush=True)


@pytest.fixture
def say(request):
    return lambda msg: emit(request.config, f"{request.node.name}: {msg}")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx():
    """One engine context for the whole GPU session (one process, one device)."""
    from pagerank_amd import Context
    c = Context(0)
    yield c
    c.close()


def cycle_graph(n: int):
    """Directed cycle 0->1->...->n-1->0 plus self-loops: vertex-transitive, PageRank exactly 1/n."""
    from pagerank_amd import HostGraph
    u = np.arange(n, dtype=np.uint32)
    return HostGraph.from_edges(n, u, (u + 1) % n, self_loops=True)


def complete_graph(n: int):
    """Complete directed graph with self-loops: PageRank exactly 1/n."""
    from pagerank_amd import HostGraph
    u, v = np.meshgrid(np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32))
    return HostGraph.from_edges(n, u.ravel(), v.ravel(), self_loops=True)


def torus_graph(rows: int, cols: int):
    """4-neighbour 2D torus (symmetric) plus self-loops: vertex-transitive."""
    from pagerank_amd import HostGraph
    i, j = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    vid = (i * cols + j).ravel()
    nbrs = [((i + 1) % rows) * cols + j, ((i - 1) % rows) * cols + j, i * cols + (j + 1) % cols,
            i * cols + (j - 1) % cols]
    src = np.concatenate([vid] * 4).astype(np.uint32)
    dst = np.concatenate([n.ravel() for n in nbrs]).astype(np.uint32)
    return HostGraph.from_edges(rows * cols, src, dst, self_loops=True)


def star_graph(n: int):
    """Hub 0 with edges to and from every other vertex, plus self-loops: one very high in-degree vertex."""
    from pagerank_amd import HostGraph
    leaves = np.arange(1, n, dtype=np.uint32)
    src = np.concatenate([leaves, np.zeros(n - 1, dtype=np.uint32)])
    dst = np.concatenate([np.zeros(n - 1, dtype=np.uint32), leaves])
    return HostGraph.from_edges(n, src, dst, self_loops=True)
