"""The C-ABI library (libudsplat.so) loads on a machine without a GPU and
exports every entry point include/udsplat.h declares, with the argument
counts the Python host layer binds.  No compute entry point is called here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "udsplat.h")


def header_prototypes():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(?:int|size_t)\s+(ude_\w+)\s*\(([^)]*)\)\s*;", src):
        args = [a.strip() for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        protos[m.group(1)] = args
    return protos


def test_header_parses():
    protos = header_prototypes()
    assert len(protos) >= 16
    assert "ude_sample" in protos and "ude_splat_backward" in protos


def test_library_exports_every_declared_symbol():
    from cudaExtensions import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_prototypes():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (ude_\w+)", out))
    assert exported == set(header_prototypes()), exported ^ set(header_prototypes())


def test_python_bindings_match_header():
    from cudaExtensions import _lib
    protos = header_prototypes()
    assert set(_lib._SIGS) == set(protos)
    for name, (res, args) in _lib._SIGS.items():
        assert len(args) == len(protos[name]), (name, len(args), len(protos[name]))
        for ca, decl in zip(args, protos[name]):
            if "*" in decl or decl.startswith("ude_stream_t"):
                assert ca is ctypes.c_void_p, (name, decl)
            elif decl.startswith("double"):
                assert ca is ctypes.c_double, (name, decl)
            elif decl.startswith("float"):
                assert ca is ctypes.c_float, (name, decl)
            elif decl.startswith("uint64_t"):
                assert ca is ctypes.c_uint64, (name, decl)
            elif decl.startswith("size_t"):
                assert ca is ctypes.c_size_t, (name, decl)
            elif decl.startswith("long "):
                assert ca is ctypes.c_long, (name, decl)
            else:
                assert decl.startswith("int ") and ca is ctypes.c_int, (name, decl)


def test_library_loads_through_host_layer():
    from cudaExtensions import _lib
    L = _lib.lib()
    assert L.ude_splat_workspace_bytes(2, 1000) >= 2 * 1000 * 64


def test_workspace_queries():
    """Host-only size queries: grow with the problem, hold the per-view tree."""
    from cudaExtensions import _lib
    L = _lib.lib()
    small = L.ude_sample_workspace_bytes(1000, 64, 64, 8)
    big = L.ude_sample_workspace_bytes(1_000_000, 512, 512, 40)
    assert 0 < small < big
    # at least the tree arrays: 2N nodes of geometry/links/aux and 289 (N - 1) coefficients
    assert big >= 1_000_000 * (2 * 64 + 4 * 289)
    assert L.ude_sample_workspace_bytes(1000, 64, 64, 64) >= L.ude_sample_workspace_bytes(1000, 64, 64, 8)
    assert L.ude_splat_workspace_bytes(1, 0) >= 256


def test_missing_library_raises(monkeypatch, tmp_path):
    from cudaExtensions import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="libudsplat.so not found"):
        _lib.lib()
