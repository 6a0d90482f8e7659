"""Graph file readers (MTX, SNAP temporal) against the reference's own readers (readMtxW, readTemporalW), compiled
from its headers into oracle/_ref, and against the formats' stated rules. CPU only."""
from __future__ import annotations

import numpy as np
import pytest

MTX_GENERAL = """%%MatrixMarket matrix coordinate real general
% a comment
%
5 5 9
1 2 0.5
2 3 1.0
3 1 2.0
3 4 1.5
4 5 1.0
5 1 3.0
1 2 0.7
2 2 1.0
4 1   0.25
"""

MTX_SYMMETRIC = """%%MatrixMarket matrix coordinate pattern symmetric
4 4 4
2 1
3 1
4 3
4 4
"""

TEMPORAL = "".join(f"{u} {v} {1000 + i}\n" for i, (u, v) in enumerate(
    [(1, 2), (2, 3), (3, 1), (1, 3), (4, 1), (2, 4), (4, 2), (1, 2), (3, 4), (4, 4), (2, 1), (3, 2)]))


def _csr(g):
    return np.asarray(g.out_off, np.int64), np.asarray(g.out_edg, np.int64)


def _expected(n, pairs):
    rows = [sorted({v for (u, v) in pairs if u == k}) for k in range(n)]
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    return off, np.array([v for r in rows for v in r], np.int64)


@pytest.mark.parametrize("text,pairs", [
    (MTX_GENERAL, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 0), (1, 1), (3, 0)]),
    (MTX_SYMMETRIC, [(1, 0), (0, 1), (2, 0), (0, 2), (3, 2), (2, 3), (3, 3)]),
    (MTX_SYMMETRIC.replace("pattern symmetric", "integer skew-symmetric"),
     [(1, 0), (0, 1), (2, 0), (0, 2), (3, 2), (2, 3), (3, 3)]),
], ids=["general", "symmetric", "skew-symmetric"])
def test_mtx_reader(tmp_path, oracle_mod, text, pairs):
    from pagerank_amd import HostGraph
    p = tmp_path / "g.mtx"
    p.write_text(text)
    g = HostGraph.from_mtx(str(p), self_loops=False)
    off, edg = _csr(g)
    eo, ee = _expected(g.N, pairs)
    assert np.array_equal(off, eo) and np.array_equal(edg, ee)
    if oracle_mod.ref_available():
        ro, re_ = oracle_mod.ref_read_mtx(p)
        assert np.array_equal(off, ro) and np.array_equal(edg, re_)


def test_mtx_reader_rejects(tmp_path):
    from pagerank_amd import HostGraph
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n3 3\n1\n2\n")
    with pytest.raises(ValueError):
        HostGraph.from_mtx(str(bad))
    out = tmp_path / "out.mtx"
    out.write_text("%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n")
    with pytest.raises(ValueError):
        HostGraph.from_mtx(str(out))


@pytest.mark.parametrize("symmetric", [False, True])
def test_temporal_reader_in_pieces(tmp_path, oracle_mod, symmetric):
    """Consecutive reads continue where the last stopped; the first k lines as a graph equal the reference's
    readTemporalW of k lines, and pieces 90% + the rest give the whole file."""
    from pagerank_amd import HostGraph, TemporalReader
    p = tmp_path / "t.txt"
    p.write_text(TEMPORAL)
    rd = TemporalReader(str(p), rows=4, symmetric=symmetric)
    s1, d1 = rd.read(9)
    s2, d2 = rd.read(100)
    s3, d3 = rd.read(5)
    rd.close()
    k = 2 if symmetric else 1
    assert len(s1) == 9 * k and len(s2) == 3 * k and len(s3) == 0
    lines = [tuple(map(int, ln.split()[:2])) for ln in TEMPORAL.splitlines()]
    first = [(u - 1, v - 1) for u, v in lines[:9]]
    want = first + ([(v, u) for u, v in first] if symmetric else [])
    assert sorted(zip(s1.tolist(), d1.tolist())) == sorted(want)
    g = HostGraph.from_edges(4, s1, d1, self_loops=False)
    if oracle_mod.ref_available():
        ro, re_ = oracle_mod.ref_read_temporal(p, 4, 9, symmetric)
        off, edg = _csr(g)
        assert np.array_equal(off, ro) and np.array_equal(edg, re_)
