"""Host-side graphs and update batches in the layout the engine takes.

A HostGraph is the prepared graph of the reference's experiment (main.cu:168-170): dense ids 0..N-1,
duplicate-free, a self-loop on every vertex, out-neighbours and in-neighbours ascending, out-degree per vertex.
Large graphs are built natively (lib/libprh_host.so, OpenMP); the arrays are zero-copy numpy views of the
native buffers, which live as long as the HostGraph.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native


def _view(ptr, n: int, dtype=np.uint32) -> np.ndarray:
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


def as_u32p(a: np.ndarray):
    return a.ctypes.data_as(_native.u32p)


class HostGraph:
    """CSR pair (out and in) of a directed graph; see the module docstring."""

    def __init__(self, handle=None, *, arrays: dict | None = None):
        self._handle = handle
        if handle is not None:
            h = _native.host()
            self.N = int(h.prhh_order(handle))
            self.M = int(h.prhh_size(handle))
            self.out_off = _view(h.prhh_out_off(handle), self.N + 1)
            self.out_edg = _view(h.prhh_out_edg(handle), self.M)
            self.in_off = _view(h.prhh_in_off(handle), self.N + 1)
            self.in_edg = _view(h.prhh_in_edg(handle), self.M)
            self.in_eid = _view(h.prhh_in_eid(handle), self.M)
            self.outdeg = _view(h.prhh_outdeg(handle), self.N)
        else:
            a = arrays
            self.N = int(len(a["out_off"]) - 1)
            self.M = int(a["out_off"][-1])
            for k in ("out_off", "out_edg", "in_off", "in_edg", "in_eid", "outdeg"):
                setattr(self, k, np.ascontiguousarray(a[k], dtype=np.uint32))

    def __del__(self):
        if getattr(self, "_handle", None) is not None:
            try:
                _native.host().prhh_free(self._handle)
            except Exception:  # interpreter teardown: the library may already be unloaded
                pass
            self._handle = None

    # --- construction -------------------------------------------------------------------------------------
    @classmethod
    def rmat(cls, scale: int, edge_factor: float = 16.0, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 42, self_loops: bool = True) -> "HostGraph":
        h = _native.host().prhh_rmat(scale, edge_factor, a, b, c, seed, 1 if self_loops else 0)
        if not h:
            raise ValueError("RMAT generation failed (scale out of range or more than 2^32 edges)")
        return cls(h)

    @classmethod
    def grid(cls, rows: int, cols: int, p: float = 0.62, seed: int = 42) -> "HostGraph":
        h = _native.host().prhh_grid(rows, cols, p, seed)
        if not h:
            raise ValueError("grid generation failed")
        return cls(h)

    @classmethod
    def from_edges(cls, n: int, src, dst, self_loops: bool = True) -> "HostGraph":
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        h = _native.host().prhh_from_edges(n, len(src), as_u32p(src), as_u32p(dst), 1 if self_loops else 0)
        if not h:
            raise ValueError("bad edge list (id >= n or n == 0)")
        return cls(h)

    @classmethod
    def from_mtx(cls, path: str, self_loops: bool = False) -> "HostGraph":
        """A Matrix Market file as a graph (readMtxW, mtx.hxx), duplicates removed; with self_loops, one self-loop
        per vertex (addSelfLoopsOmp, as main.cu:71 does before solving)."""
        src, dst, order = read_mtx_edges(path)
        return cls.from_edges(order, src, dst, self_loops=self_loops)

    @classmethod
    def from_csr(cls, out_off, out_edg) -> "HostGraph":
        """From an out-CSR given as numpy arrays (rows ascending, no self-loops added). Builds the in-CSR."""
        out_off = np.ascontiguousarray(out_off, dtype=np.uint32)
        out_edg = np.ascontiguousarray(out_edg, dtype=np.uint32)
        n = len(out_off) - 1
        src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(out_off.astype(np.int64)))
        order = np.lexsort((src, out_edg))
        in_edg = src[order]
        in_eid = order.astype(np.uint32)
        counts = np.bincount(out_edg, minlength=n)
        in_off = np.zeros(n + 1, dtype=np.uint32)
        np.cumsum(counts, out=in_off[1:])
        outdeg = np.diff(out_off.astype(np.int64)).astype(np.uint32)
        return cls(arrays=dict(out_off=out_off, out_edg=out_edg, in_off=in_off, in_edg=in_edg, in_eid=in_eid,
                               outdeg=outdeg))

    def _require_handle(self):
        """The native handle; a numpy-built graph is re-homed into the native library first."""
        if self._handle is None:
            src = self.out_sources()
            h = _native.host().prhh_from_edges(self.N, len(src), as_u32p(src), as_u32p(self.out_edg), 0)
            if not h:
                raise ValueError("cannot re-home graph")
            g = HostGraph(h)
            for k in ("out_off", "out_edg", "in_off", "in_edg", "in_eid", "outdeg"):
                setattr(self, k, getattr(g, k))
            self._handle, g._handle = h, None
        return self._handle

    # --- views ----------------------------------------------------------------------------------------------
    def out_sources(self) -> np.ndarray:
        return np.repeat(np.arange(self.N, dtype=np.uint32), np.diff(self.out_off.astype(np.int64)))

    def indegree(self) -> np.ndarray:
        return np.diff(self.in_off.astype(np.int64))

    def dense_transition(self) -> np.ndarray:
        """Dense column-stochastic matrix A with A[v,u] = 1/outdeg(u) for each edge u->v (small graphs only)."""
        A = np.zeros((self.N, self.N))
        src = self.out_sources()
        A[self.out_edg, src] = 1.0 / self.outdeg[src]
        return A


def _take_edges(eh) -> tuple[np.ndarray, np.ndarray, int]:
    h = _native.host()
    n = int(h.prhh_edges_count(eh))
    src = _view(C.cast(h.prhh_edges_src(eh), _native.u32p), n).copy()
    dst = _view(C.cast(h.prhh_edges_dst(eh), _native.u32p), n).copy()
    order = int(h.prhh_edges_order(eh))
    h.prhh_edges_free(eh)
    return src, dst, order


def read_mtx_edges(path: str) -> tuple[np.ndarray, np.ndarray, int]:
    """Matrix Market coordinate file -> (src, dst, order) with 0-based ids (file vertex k -> k - 1), mirrors added
    for symmetric files (readMtxDo, mtx.hxx:120-140)."""
    eh = _native.host().prhh_read_mtx(str(path).encode())
    if not eh:
        raise ValueError(f"cannot read {path} as a Matrix Market coordinate graph")
    return _take_edges(eh)


class TemporalReader:
    """SNAP temporal edge file read in consecutive pieces (readTemporalDo, snap.hxx:40-60): read(n) returns the
    next n lines as 0-based (src, dst), with mirrors when symmetric."""

    def __init__(self, path: str, rows: int, symmetric: bool = False):
        self._h = _native.host().prhh_temporal_open(str(path).encode())
        if not self._h:
            raise ValueError(f"cannot open {path}")
        self.rows, self.symmetric = rows, symmetric

    def read(self, lines: int) -> tuple[np.ndarray, np.ndarray]:
        eh = _native.host().prhh_temporal_read(self._h, self.rows, lines, 1 if self.symmetric else 0)
        if not eh:
            raise ValueError("vertex id outside 1..rows in the temporal file")
        src, dst, _ = _take_edges(eh)
        return src, dst

    def close(self):
        if getattr(self, "_h", None):
            _native.host().prhh_temporal_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Batch:
    """A tidied update batch: deletions (del_u[i], del_v[i]) and insertions (ins_u[i], ins_v[i]), each sorted by
    (source, target) and unique; deletions exist in x, insertions do not (tidyBatchUpdateU, batch.hxx:201-208)."""
    del_u: np.ndarray
    del_v: np.ndarray
    ins_u: np.ndarray
    ins_v: np.ndarray

    @property
    def deletions(self) -> int:
        return len(self.del_u)

    @property
    def insertions(self) -> int:
        return len(self.ins_u)

    @classmethod
    def _from_handle(cls, bh) -> "Batch":
        h = _native.host()
        nd, ni = int(h.prhh_batch_deletions(bh)), int(h.prhh_batch_insertions(bh))
        b = cls(_view(h.prhh_batch_del_u(bh), nd).copy(), _view(h.prhh_batch_del_v(bh), nd).copy(),
                _view(h.prhh_batch_ins_u(bh), ni).copy(), _view(h.prhh_batch_ins_v(bh), ni).copy())
        h.prhh_batch_free(bh)
        return b

    @classmethod
    def random(cls, x: HostGraph, size: int, insert_fraction: float = 0.8, seed: int = 1) -> "Batch":
        bh = _native.host().prhh_random_batch(x._require_handle(), size, insert_fraction, seed)
        if not bh:
            raise ValueError("batch generation failed")
        return cls._from_handle(bh)

    @classmethod
    def tidy(cls, x: HostGraph, deletions, insertions) -> "Batch":
        d = np.asarray(deletions, dtype=np.uint32).reshape(-1, 2)
        i = np.asarray(insertions, dtype=np.uint32).reshape(-1, 2)
        du, dv = np.ascontiguousarray(d[:, 0]), np.ascontiguousarray(d[:, 1])
        iu, iv = np.ascontiguousarray(i[:, 0]), np.ascontiguousarray(i[:, 1])
        bh = _native.host().prhh_make_batch(x._require_handle(), as_u32p(du), as_u32p(dv), len(du), as_u32p(iu),
                                            as_u32p(iv), len(iu))
        if not bh:
            raise ValueError("bad batch (insertion id out of range)")
        return cls._from_handle(bh)

    @classmethod
    def empty(cls) -> "Batch":
        z = np.zeros(0, dtype=np.uint32)
        return cls(z, z.copy(), z.copy(), z.copy())


def apply_batch(x: HostGraph, b: Batch) -> HostGraph:
    """y = x with b applied (deletions removed, insertions added; applyBatchUpdateOmpU, batch.hxx:240-247)."""
    h = _native.host()
    bh = h.prhh_make_batch(x._require_handle(), as_u32p(b.del_u), as_u32p(b.del_v), len(b.del_u),
                           as_u32p(b.ins_u), as_u32p(b.ins_v), len(b.ins_u))
    if not bh:
        raise ValueError("bad batch")
    yh = h.prhh_apply_batch(x._require_handle(), bh)
    h.prhh_batch_free(bh)
    if not yh:
        raise ValueError("batch application failed")
    return HostGraph(yh)

