"""prh_partition_noalloc (device arrays, fixed scratch, no device allocation) against prh_partition, through the
C ABI: bit-exact order and partition point on two calls in a row, the second on a smaller graph."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device_u32(torch, a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def _host_u32(t) -> np.ndarray:
    return t.cpu().numpy().view(np.uint32)


def _check_one(torch, ctx, lib, offsets: np.ndarray, d: int, oracle_mod):
    n = len(offsets) - 1
    ref_order, ref_np = ctx.partition(offsets, d)  # prh_partition: host arrays, context scratch
    off = _device_u32(torch, offsets)
    guard = 64
    order = torch.full((n + guard,), -1, dtype=torch.int32, device="cuda")
    point = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    used = ctx.memory_used()
    st = lib.prh_partition_noalloc(ctx.handle, n, C.c_void_p(off.data_ptr()), d, C.c_void_p(order.data_ptr()),
                                   C.c_void_p(point.data_ptr()))
    assert st == 0, lib.prh_context_last_error(ctx.handle)
    torch.cuda.synchronize()  # the call is asynchronous on the context's stream
    assert ctx.memory_used() == used  # nothing allocated through the context
    got = _host_u32(order)
    assert np.array_equal(got[:n], ref_order)
    assert np.all(got[n:] == 0xFFFFFFFF)  # nothing written past N
    assert int(_host_u32(point)[0]) == ref_np
    o_order, o_np = oracle_mod.partition(offsets, d)
    assert np.array_equal(ref_order, o_order) and ref_np == o_np


def test_noalloc_partition_matches_prh_partition_twice(ctx, oracle_mod):
    import torch
    from pagerank_amd import HostGraph, _native
    lib = _native.engine()
    big = HostGraph.rmat(14, 16, seed=7)
    small = HostGraph.rmat(10, 8, seed=3)
    print(f"[noalloc] partition of RMAT-14 in-offsets (N={big.N}), then RMAT-10 (N={small.N})", flush=True)
    _check_one(torch, ctx, lib, big.in_off, 64, oracle_mod)
    _check_one(torch, ctx, lib, small.in_off, 8, oracle_mod)


def test_noalloc_partition_edges(ctx):
    """Empty graph: partition point 0; above the fixed scratch's limit: a clear error, nothing launched."""
    import torch
    from pagerank_amd import _native
    lib = _native.engine()
    point = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    buf = torch.zeros(16, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert lib.prh_partition_noalloc(ctx.handle, 0, None, 64, None, C.c_void_p(point.data_ptr())) == 0
    torch.cuda.synchronize()
    assert int(point.item()) == 0
    p = C.c_void_p(buf.data_ptr())
    st = lib.prh_partition_noalloc(ctx.handle, (1 << 28) + 1, p, 64, p, p)
    assert st == 1
    assert b"exceed its fixed scratch" in lib.prh_context_last_error(ctx.handle)
