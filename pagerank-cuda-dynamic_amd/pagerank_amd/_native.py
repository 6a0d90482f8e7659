"""Loading of the engine's native libraries (built in-tree by the package Makefile).

There is no fallback: if ``lib/libpagerank_hip.so`` is missing the import of the engine fails with an error
that says how to build it. The host helper library ``lib/libprh_host.so`` (graph generators) is separate and
needs no GPU.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent          # .../pagerank-cuda-dynamic_amd
REPO_DIR = PKG_DIR.parent
LIB_DIR = PKG_DIR / "lib"
ENGINE_LIB = LIB_DIR / "libpagerank_hip.so"
HOST_LIB = LIB_DIR / "libprh_host.so"

u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
vp = C.c_void_p


def build(force: bool = True, jobs: int = 8) -> None:
    """Compile both libraries with the package Makefile (hipcc for gfx950, g++ for the host helpers)."""
    cmd = ["make", "-C", str(PKG_DIR), f"-j{jobs}", "all"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)


class prh_options(C.Structure):
    _fields_ = [("repeat", C.c_int), ("tolerance", C.c_double), ("frontier_tolerance", C.c_double),
                ("prune_tolerance", C.c_double), ("damping", C.c_double), ("max_iterations", C.c_int)]


class prh_stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("time", C.c_float), ("initialization_time", C.c_float),
                ("marking_time", C.c_float), ("computation_time", C.c_float), ("preparation_time", C.c_float)]


class prh_tuning(C.Structure):
    _fields_ = [("switch_degree", C.c_uint32), ("huge_degree", C.c_uint32), ("low_width", C.c_int),
                ("check_interval", C.c_int), ("source_slices", C.c_int), ("block_edges", C.c_uint32),
                ("extension_direction", C.c_int), ("blocking", C.c_int), ("layout", C.c_int),
                ("window", C.c_int)]


# Every symbol include/pagerank_hip.h declares, with its ctypes signature (restype, argtypes).
ENGINE_SIGNATURES = {
    "prh_default_options": (None, [C.POINTER(prh_options)]),
    "prh_default_tuning": (None, [C.POINTER(prh_tuning)]),
    "prh_context_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "prh_context_destroy": (None, [vp]),
    "prh_context_last_error": (C.c_char_p, [vp]),
    "prh_context_set_memory_limit": (C.c_int, [vp, C.c_size_t]),
    "prh_context_memory_used": (C.c_size_t, [vp]),
    "prh_context_set_tuning": (C.c_int, [vp, C.POINTER(prh_tuning)]),
    "prh_context_get_tuning": (C.c_int, [vp, C.POINTER(prh_tuning)]),
    "prh_context_stream": (vp, [vp]),
    "prh_graph_create": (C.c_int, [vp, C.c_uint32, C.c_uint64, u32p, u32p, u32p, u32p, u32p, C.POINTER(vp)]),
    "prh_graph_destroy": (None, [vp]),
    "prh_graph_order": (C.c_uint32, [vp]),
    "prh_graph_size": (C.c_uint64, [vp]),
    "prh_graph_update": (C.c_int, [vp, vp, u32p, u32p, C.c_size_t, u32p, u32p, C.c_size_t, C.POINTER(C.c_float)]),
    "prh_graph_csr": (C.c_int, [vp, vp, C.c_int, u32p, u32p]),
    "prh_solve": (C.c_int, [vp, vp, C.c_int, C.c_int, C.POINTER(prh_options), u32p, u32p, C.c_size_t, u32p,
                            C.c_size_t, f64p, f64p, C.POINTER(prh_stats)]),
    "prh_solve_device": (C.c_int, [vp, vp, C.c_int, C.c_int, C.POINTER(prh_options), vp, vp, C.c_size_t, vp,
                                   C.c_size_t, vp, vp, C.POINTER(prh_stats)]),
    "prh_partition": (C.c_int, [vp, C.c_uint32, u32p, C.c_uint32, u32p, u32p]),
    "prh_partition_noalloc": (C.c_int, [vp, C.c_uint32, vp, C.c_uint32, vp, vp]),
    "prh_mark_affected": (C.c_int, [vp, vp, C.c_int, u32p, u32p, C.c_size_t, u32p, C.c_size_t, u8p]),
    "prh_ppr_batched": (C.c_int, [vp, vp, C.c_int, u32p, u32p, C.POINTER(prh_options), f64p, C.POINTER(C.c_int),
                                  C.POINTER(prh_stats)]),
    "prh_weighted_pagerank": (C.c_int, [vp, vp, vp, vp, C.POINTER(prh_options), vp, C.POINTER(prh_stats), C.c_int]),
    "prh_weighted_adjoint": (C.c_int, [vp, vp, vp, vp, vp, C.POINTER(prh_options), vp, vp, vp,
                                       C.POINTER(prh_stats), C.c_int]),
    "prh_autotune": (C.c_int, [vp, vp, C.c_char_p, C.c_int, C.POINTER(prh_tuning), C.POINTER(C.c_double),
                              C.POINTER(C.c_int)]),
    "prh_stream_create": (C.c_int, [vp, vp, C.c_int, C.POINTER(prh_options), C.c_uint32, C.c_uint64, f64p,
                                   C.POINTER(vp)]),
    "prh_stream_update": (C.c_int, [vp, u32p, u32p, C.c_size_t, u32p, u32p, C.c_size_t, f64p, C.POINTER(prh_stats)]),
    "prh_stream_ranks": (C.c_int, [vp, f64p]),
    "prh_stream_size": (C.c_uint64, [vp]),
    "prh_stream_csr": (C.c_int, [vp, C.c_int, u32p, u32p]),
    "prh_stream_destroy": (None, [vp]),
    "prh_kernel_timing_enable": (C.c_int, [vp, C.c_int]),
    "prh_kernel_timing_read": (C.c_int, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
}

HOST_SIGNATURES = {
    "prhh_rmat": (vp, [C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int]),
    "prhh_grid": (vp, [C.c_uint32, C.c_uint32, C.c_double, C.c_uint64]),
    "prhh_from_edges": (vp, [C.c_uint32, C.c_uint64, u32p, u32p, C.c_int]),
    "prhh_read_mtx": (vp, [C.c_char_p]),
    "prhh_temporal_open": (vp, [C.c_char_p]),
    "prhh_temporal_read": (vp, [vp, C.c_uint64, C.c_uint64, C.c_int]),
    "prhh_temporal_close": (None, [vp]),
    "prhh_edges_count": (C.c_uint64, [vp]),
    "prhh_edges_order": (C.c_uint64, [vp]),
    "prhh_edges_symmetric": (C.c_int, [vp]),
    "prhh_edges_src": (vp, [vp]),
    "prhh_edges_dst": (vp, [vp]),
    "prhh_edges_free": (None, [vp]),
    "prhh_free": (None, [vp]),
    "prhh_order": (C.c_uint32, [vp]),
    "prhh_size": (C.c_uint64, [vp]),
    "prhh_out_off": (u32p, [vp]),
    "prhh_out_edg": (u32p, [vp]),
    "prhh_in_off": (u32p, [vp]),
    "prhh_in_edg": (u32p, [vp]),
    "prhh_in_eid": (u32p, [vp]),
    "prhh_outdeg": (u32p, [vp]),
    "prhh_random_batch": (vp, [vp, C.c_uint64, C.c_double, C.c_uint64]),
    "prhh_make_batch": (vp, [vp, u32p, u32p, C.c_uint64, u32p, u32p, C.c_uint64]),
    "prhh_batch_free": (None, [vp]),
    "prhh_batch_deletions": (C.c_uint64, [vp]),
    "prhh_batch_insertions": (C.c_uint64, [vp]),
    "prhh_batch_del_u": (u32p, [vp]),
    "prhh_batch_del_v": (u32p, [vp]),
    "prhh_batch_ins_u": (u32p, [vp]),
    "prhh_batch_ins_v": (u32p, [vp]),
    "prhh_apply_batch": (vp, [vp, vp]),
}


def _load(path: Path, signatures: dict, what: str) -> C.CDLL:
    if not path.exists():
        raise RuntimeError(
            f"{what} not built: {path} is missing. Build it with `make -C {PKG_DIR}` "
            f"(or __graft_entry__.build()); there is no fallback path.")
    lib = C.CDLL(str(path))
    for name, (res, args) in signatures.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_engine = None
_host = None


def engine() -> C.CDLL:
    """The HIP engine library (kernels + C ABI)."""
    global _engine
    if _engine is None:
        _engine = _load(ENGINE_LIB, ENGINE_SIGNATURES, "HIP engine library")
    return _engine


def host() -> C.CDLL:
    """The host-side graph helper library."""
    global _host
    if _host is None:
        _host = _load(HOST_LIB, HOST_SIGNATURES, "host graph library")
    return _host


def env_int(name: str, default: int) -> int:
    try:
        return int(os.environ.get(name, default))
    except ValueError:
        return default
