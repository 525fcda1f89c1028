"""The C-ABI library loads and exports every entry point include/qcmd.h
declares (no compute calls: this runs without a GPU)."""

import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "qcmd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(qc_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "qc_eri" in names and "qc_symv" in names and len(names) >= 10


def test_library_exports_every_declared_symbol():
    from jax_qc import _native

    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libqcmd.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    bound = set(_native.exported_symbols())
    assert set(declared()) <= bound | {"qc_last_error"}, set(declared()) - bound


def test_product_path_fails_loudly_without_library(tmp_path, monkeypatch):
    from jax_qc import _native

    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _native.lib()
