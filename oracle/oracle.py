"""ctypes front of the oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/_build/libprh_oracle.so (the CPU restatement in prh_oracle.c) and, when it has been built,
oracle/_ref/libprh_ref.so (the reference's own CPU path compiled from its headers by oracle/Makefile).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the checker.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "_build" / "libprh_oracle.so"
REF_LIB = HERE / "_ref" / "libprh_ref.so"

u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int)

STATIC, NAIVE_DYNAMIC, DYNAMIC_TRAVERSAL, DYNAMIC_FRONTIER, PRUNE_DYNAMIC_FRONTIER = range(5)

_SIG = {
    "orc_partition": (C.c_uint32, [C.c_uint32, u32p, C.c_uint32, u32p]),
    "orc_mark_frontier": (None, [C.c_uint32, u32p, u32p, u32p, u32p, C.c_size_t, u32p, C.c_size_t, u8p, u8p]),
    "orc_mark_traversal": (None, [C.c_uint32, u32p, u32p, u32p, u32p, C.c_size_t, u32p, C.c_size_t, u8p]),
    "orc_pagerank_cuda": (C.c_int, [C.c_int, C.c_uint32, u32p, u32p, u32p, u32p, u32p, u32p, u32p, C.c_size_t,
                                    u32p, C.c_size_t, f64p, C.c_double, C.c_double, C.c_int, C.c_double,
                                    C.c_double, f64p, ip, u8p]),
    "orc_static_omp": (C.c_int, [C.c_uint32, u32p, u32p, u32p, C.c_double, C.c_double, C.c_int, C.c_int, f64p, ip,
                                 f64p]),
    "orc_pagerank_async": (C.c_int, [C.c_int, C.c_uint32, u32p, u32p, u32p, u32p, u32p, u32p, u32p, C.c_size_t,
                                     u32p, C.c_size_t, f64p, C.c_double, C.c_double, C.c_int, C.c_double,
                                     C.c_double, C.c_int, f64p, ip, f64p]),
    "orc_power_iteration": (C.c_int, [C.c_uint32, u32p, u32p, u32p, f64p, C.c_double, C.c_int, f64p, ip]),
    "orc_ppr": (C.c_int, [C.c_uint32, u32p, u32p, u32p, f64p, C.c_double, C.c_double, C.c_int, f64p, ip]),
    "orc_weighted_forward": (C.c_int, [C.c_uint32, u32p, u32p, f32p, f64p, f64p, C.c_double, C.c_double, C.c_int,
                                       f64p, ip]),
    "orc_weighted_adjoint": (C.c_int, [C.c_uint32, u32p, u32p, f32p, f64p, f64p, C.c_double, C.c_double, C.c_int,
                                       f64p, ip]),
    "orc_weighted_grad": (None, [C.c_uint32, u32p, u32p, f64p, f64p, f64p, f64p, C.c_double, f64p, f64p]),
}
_REF_SIG = {
    "ref_pagerank": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint32, u32p, u32p, u32p, u32p, C.c_size_t, u32p,
                               u32p, C.c_size_t, f64p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                               C.c_int, f64p, ip, C.POINTER(C.c_float)]),
    "ref_affected": (C.c_int, [C.c_int, C.c_uint32, u32p, u32p, u32p, u32p, C.c_size_t, u32p, u32p, C.c_size_t,
                               u8p]),
    "ref_apply_batch": (C.c_int64, [C.c_uint32, u32p, u32p, u32p, u32p, C.c_size_t, u32p, u32p, C.c_size_t, u32p,
                                    u32p]),
    "ref_read_mtx": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), u32p, u32p]),
    "ref_read_temporal": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint64, C.c_int, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint64), u32p, u32p]),
}

_lib = None
_ref = None


def build(reference: bool = False) -> None:
    subprocess.run(["make", "-C", str(HERE), "-B", "all"], check=True)
    if reference:
        subprocess.run(["make", "-C", str(HERE), "-B", "ref"], check=True)


def _load(path: Path, sig: dict) -> C.CDLL:
    lib = C.CDLL(str(path))
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype, f.argtypes = r, a
    return lib


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            build()
        _lib = _load(ORACLE_LIB, _SIG)
    return _lib


def ref_available() -> bool:
    return REF_LIB.exists()


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        _ref = _load(REF_LIB, _REF_SIG)
    return _ref


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, (a.ctypes.data_as(u32p) if len(a) else None)


def partition(offsets, d: int = 64):
    off, p = _u32(offsets)
    n = len(off) - 1
    out = np.zeros(n, dtype=np.uint32)
    np_ = lib().orc_partition(n, p, d, out.ctypes.data_as(u32p))
    return out, int(np_)


def _batch(batch):
    if batch is None:
        z = np.zeros(0, dtype=np.uint32)
        return z, z, z, z
    return (np.ascontiguousarray(batch.del_u, np.uint32), np.ascontiguousarray(batch.del_v, np.uint32),
            np.ascontiguousarray(batch.ins_u, np.uint32), np.ascontiguousarray(batch.ins_v, np.uint32))


def pagerank_cuda(method, y, batch=None, q=None, damping=0.85, tolerance=1e-10, max_iterations=500,
                  frontier_tolerance=1e-6, prune_tolerance=1e-6):
    """Sequential restatement of the reference's GPU path on graph y (a HostGraph). Returns
    (ranks, iterations, affected flags after marking)."""
    du, dv, iu, _ = _batch(batch)
    N = y.N
    ranks = np.zeros(N)
    it = C.c_int()
    vaff = np.zeros(N, dtype=np.uint8)
    qq = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
    st = lib().orc_pagerank_cuda(
        method, N, _u32(y.in_off)[1], _u32(y.in_edg)[1], _u32(y.outdeg)[1], _u32(y.out_off)[1], _u32(y.out_edg)[1],
        _u32(du)[1], _u32(dv)[1], len(du), _u32(iu)[1], len(iu),
        qq.ctypes.data_as(f64p) if qq is not None else None, damping, tolerance, max_iterations,
        frontier_tolerance, prune_tolerance, ranks.ctypes.data_as(f64p), C.byref(it), vaff.ctypes.data_as(u8p))
    if st != 0:
        raise ValueError("oracle rejected arguments")
    return ranks, int(it.value), vaff


def mark(method, y, batch):
    du, dv, iu, _ = _batch(batch)
    vaff = np.zeros(y.N, dtype=np.uint8)
    naff = np.zeros(y.N, dtype=np.uint8)
    args = (y.N, _u32(y.out_off)[1], _u32(y.out_edg)[1], _u32(du)[1], _u32(dv)[1], len(du), _u32(iu)[1], len(iu))
    if method == DYNAMIC_TRAVERSAL:
        lib().orc_mark_traversal(*args, vaff.ctypes.data_as(u8p))
    else:
        lib().orc_mark_frontier(*args, vaff.ctypes.data_as(u8p), naff.ctypes.data_as(u8p))
    return vaff


def static_omp(y, damping=0.85, tolerance=1e-10, max_iterations=500, threads=0):
    ranks = np.zeros(y.N)
    it = C.c_int()
    ms = C.c_double()
    lib().orc_static_omp(y.N, _u32(y.in_off)[1], _u32(y.in_edg)[1], _u32(y.outdeg)[1], damping, tolerance,
                         max_iterations, threads, ranks.ctypes.data_as(f64p), C.byref(it), C.byref(ms))
    return ranks, int(it.value), float(ms.value)


def pagerank_async(method, y, batch=None, q=None, threads=0, damping=0.85, tolerance=1e-10, max_iterations=500,
                   frontier_tolerance=1e-6, prune_tolerance=1e-6):
    """The reference's asynchronous OpenMP solver `method` on the updated graph y (a HostGraph; batch = the tidied
    batch that produced it, q = previous ranks). Returns (ranks, iterations, ms)."""
    du, dv, iu, _ = _batch(batch)
    ranks = np.zeros(y.N)
    it = C.c_int()
    ms = C.c_double()
    qq = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
    st = lib().orc_pagerank_async(
        method, y.N, _u32(y.in_off)[1], _u32(y.in_edg)[1], _u32(y.outdeg)[1], _u32(y.out_off)[1], _u32(y.out_edg)[1],
        _u32(du)[1], _u32(dv)[1], len(du), _u32(iu)[1], len(iu), qq.ctypes.data_as(f64p) if qq is not None else None,
        damping, tolerance, max_iterations, frontier_tolerance, prune_tolerance, threads, ranks.ctypes.data_as(f64p),
        C.byref(it), C.byref(ms))
    if st != 0:
        raise ValueError("oracle rejected arguments")
    return ranks, int(it.value), float(ms.value)


def power_iteration(y, s=None, damping=0.85, max_iterations=10000):
    ranks = np.zeros(y.N)
    it = C.c_int()
    ss = None if s is None else np.ascontiguousarray(s, dtype=np.float64)
    lib().orc_power_iteration(y.N, _u32(y.in_off)[1], _u32(y.in_edg)[1], _u32(y.outdeg)[1],
                              ss.ctypes.data_as(f64p) if ss is not None else None, damping, max_iterations,
                              ranks.ctypes.data_as(f64p), C.byref(it))
    return ranks, int(it.value)


def ppr(y, s, damping=0.85, tolerance=1e-10, max_iterations=500):
    ranks = np.zeros(y.N)
    it = C.c_int()
    ss = np.ascontiguousarray(s, dtype=np.float64)
    lib().orc_ppr(y.N, _u32(y.in_off)[1], _u32(y.in_edg)[1], _u32(y.outdeg)[1], ss.ctypes.data_as(f64p), damping,
                  tolerance, max_iterations, ranks.ctypes.data_as(f64p), C.byref(it))
    return ranks, int(it.value)


def weighted_forward(y, w_out, s, damping=0.85, tolerance=1e-10, max_iterations=500):
    """Weighted forward solve; w_out is aligned with y.out_edg (FP32)."""
    w_out = np.ascontiguousarray(w_out, dtype=np.float32)
    w_in = np.ascontiguousarray(w_out[y.in_eid])
    W = np.add.reduceat(w_out.astype(np.float64), y.out_off[:-1].astype(np.int64)) if y.M else np.zeros(y.N)
    ranks = np.zeros(y.N)
    it = C.c_int()
    ss = np.ascontiguousarray(s, dtype=np.float64)
    lib().orc_weighted_forward(y.N, _u32(y.in_off)[1], _u32(y.in_edg)[1], w_in.ctypes.data_as(f32p),
                               W.ctypes.data_as(f64p), ss.ctypes.data_as(f64p), damping, tolerance, max_iterations,
                               ranks.ctypes.data_as(f64p), C.byref(it))
    return ranks, int(it.value), W


def weighted_adjoint(y, w_out, W, g, damping=0.85, tolerance=1e-10, max_iterations=500):
    w_out = np.ascontiguousarray(w_out, dtype=np.float32)
    lam = np.zeros(y.N)
    it = C.c_int()
    gg = np.ascontiguousarray(g, dtype=np.float64)
    W = np.ascontiguousarray(W, dtype=np.float64)
    lib().orc_weighted_adjoint(y.N, _u32(y.out_off)[1], _u32(y.out_edg)[1], w_out.ctypes.data_as(f32p),
                               W.ctypes.data_as(f64p), gg.ctypes.data_as(f64p), damping, tolerance, max_iterations,
                               lam.ctypes.data_as(f64p), C.byref(it))
    return lam, int(it.value)


def weighted_grad(y, W, r, lam, g, damping=0.85):
    gw = np.zeros(y.M)
    gs = np.zeros(y.N)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (W, r, lam, g)]
    lib().orc_weighted_grad(y.N, _u32(y.out_off)[1], _u32(y.out_edg)[1], *[a.ctypes.data_as(f64p) for a in arrs],
                            damping, gw.ctypes.data_as(f64p), gs.ctypes.data_as(f64p))
    return gw, gs


# --- the reference's own CPU path (oracle/_ref) ---------------------------------------------------------------
def ref_pagerank(method, x, batch=None, q=None, asynchronous=False, threads=0, damping=0.85, tolerance=1e-10,
                 max_iterations=500, frontier_tolerance=1e-6, prune_tolerance=1e-6, repeat=1):
    """Run the reference's CPU solver on y = x + batch (x given as a HostGraph). Returns (ranks, iters, ms)."""
    du, dv, iu, iv = _batch(batch)
    ranks = np.zeros(x.N)
    it = C.c_int()
    ms = C.c_float()
    qq = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
    st = ref().ref_pagerank(method, 1 if asynchronous else 0, threads, x.N, _u32(x.out_off)[1], _u32(x.out_edg)[1],
                            _u32(du)[1], _u32(dv)[1], len(du), _u32(iu)[1], _u32(iv)[1], len(iu),
                            qq.ctypes.data_as(f64p) if qq is not None else None, tolerance, frontier_tolerance,
                            prune_tolerance, damping, max_iterations, repeat, ranks.ctypes.data_as(f64p),
                            C.byref(it), C.byref(ms))
    if st != 0:
        raise ValueError("reference rejected arguments")
    return ranks, int(it.value), float(ms.value)


def ref_affected(kind, x, batch):
    du, dv, iu, iv = _batch(batch)
    out = np.zeros(x.N, dtype=np.uint8)
    ref().ref_affected(kind, x.N, _u32(x.out_off)[1], _u32(x.out_edg)[1], _u32(du)[1], _u32(dv)[1], len(du),
                       _u32(iu)[1], _u32(iv)[1], len(iu), out.ctypes.data_as(u8p))
    return out


def ref_apply_batch(x, batch):
    du, dv, iu, iv = _batch(batch)
    y_off = np.zeros(x.N + 1, dtype=np.uint32)
    y_edg = np.zeros(x.M + len(iu) + 1, dtype=np.uint32)
    m = ref().ref_apply_batch(x.N, _u32(x.out_off)[1], _u32(x.out_edg)[1], _u32(du)[1], _u32(dv)[1], len(du),
                              _u32(iu)[1], _u32(iv)[1], len(iu), y_off.ctypes.data_as(u32p),
                              y_edg.ctypes.data_as(u32p))
    return y_off, y_edg[:m]


def _ref_csr(call):
    n, m = C.c_uint32(), C.c_uint64()
    if call(C.byref(n), C.byref(m), None, None) != 0:
        raise ValueError("reference reader failed")
    off = np.zeros(n.value + 1, dtype=np.uint32)
    edg = np.zeros(max(m.value, 1), dtype=np.uint32)
    call(C.byref(n), C.byref(m), _u32(off)[1], _u32(edg)[1])
    return off, edg[:m.value]


def ref_read_mtx(path):
    """The reference's readMtxW (mtx.hxx) as an out-CSR over the existing vertices in key order."""
    p = str(path).encode()
    return _ref_csr(lambda n, m, o, e: ref().ref_read_mtx(p, n, m, o, e))


def ref_read_temporal(path, rows, lines, symmetric=False):
    """The reference's readTemporalW (snap.hxx) on the first `lines` lines, as ref_read_mtx."""
    p = str(path).encode()
    return _ref_csr(lambda n, m, o, e: ref().ref_read_temporal(p, rows, lines, 1 if symmetric else 0, n, m, o, e))

