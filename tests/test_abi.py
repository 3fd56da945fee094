"""The C-ABI library loads and exports every symbol include/zandhip.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from zandpack_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "zandhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(zh_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_header_symbols():
    if not os.path.exists(_native._LIB_PATH):
        from zandpack_amd import build
        build.build()
    lib = ctypes.CDLL(_native._LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    for nm in names:
        assert hasattr(lib, nm), nm
    # the Python binding declares exactly the header's entry points
    assert sorted(_native.SIGNATURES) == names


def test_version_and_error_channel():
    L = _native.lib()
    assert L.zh_abi_version() == 1
    # argument validation happens before any device call
    d = _native.zh_dims(0, 4, 2, 3, 1, 0, 0.658, 0)
    h = ctypes.c_void_p()
    rc = L.zh_problem_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == -1
    assert b"positive" in L.zh_last_error()


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(_native, "_LIB_PATH", "/nonexistent/libzandhip.so")
    try:
        _native.lib()
    except ImportError as e:
        assert "no CPU fallback" in str(e)
    else:
        raise AssertionError("expected ImportError")
