"""The C-ABI libraries load and export every symbol their headers declare (CPU only: no compute calls)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent


def declared(header: Path) -> list[str]:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prh_[a-z0-9_]+)\s*\(", text)))


def test_engine_exports_every_declared_symbol():
    from pagerank_amd import _native
    lib = _native.engine()
    names = declared(REPO / "include" / "pagerank_hip.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds exactly these
    assert set(names) == set(_native.ENGINE_SIGNATURES)


def test_default_options_match_reference():
    """prh_default_options gives PagerankOptions' defaults (pagerank.hxx:52)."""
    from pagerank_amd import PagerankOptions, _native
    o = _native.prh_options()
    _native.engine().prh_default_options(C.byref(o))
    p = PagerankOptions()
    assert (o.repeat, o.tolerance, o.frontier_tolerance, o.prune_tolerance, o.damping, o.max_iterations) == \
        (p.repeat, p.tolerance, p.frontierTolerance, p.pruneTolerance, p.damping, p.maxIterations) == \
        (1, 1e-10, 1e-6, 1e-6, 0.85, 500)


def test_no_device_gives_status_not_exit():
    """Without a usable device the context call returns a status (the reference would exit the process)."""
    from pagerank_amd import _native
    import os
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert _native.engine().prh_context_create(0, C.byref(h)) == 4


def test_host_library_exports():
    from pagerank_amd import _native
    lib = _native.host()
    for n in _native.HOST_SIGNATURES:
        assert hasattr(lib, n)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from pagerank_amd import _native
    monkeypatch.setattr(_native, "_engine", None)
    monkeypatch.setattr(_native, "ENGINE_LIB", tmp_path / "absent.so")
    with pytest.raises(RuntimeError, match="not built"):
        _native.engine()


def test_default_tuning_matches_python_mirror():
    """prh_default_tuning fills exactly the fields the ctypes mirror declares (the reference's fixed launch constants
    where it has them, the banded plan off) and writes nothing past them."""
    from pagerank_amd import Tuning, _native

    class Guarded(C.Structure):
        _fields_ = [("t", _native.prh_tuning), ("guard", C.c_uint32 * 8)]

    g = Guarded()
    for i in range(8):
        g.guard[i] = 0xA5A5A5A5
    _native.engine().prh_default_tuning(C.byref(g.t))
    assert list(g.guard) == [0xA5A5A5A5] * 8
    got = {name: getattr(g.t, name) for name, _ in _native.prh_tuning._fields_}
    assert got == Tuning().__dict__
    assert (got["switch_degree"], got["blocking"], got["layout"], got["window"]) == (64, 0, 0, 0)
