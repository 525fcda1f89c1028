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


def _tile_pass_blocks(T, seg, nrhs, grouped):
    """64-double blocks of partial sums one tile pass writes, counted by
    enumerating its work items: per tile row R, R // seg + 1 segments, one
    row block each and one column block per tile; grouped (one right-hand
    side, rows in fours): per group, (last row) // seg + 1 work-groups of
    four row blocks each and one column block per column up to its last row."""
    if not grouped:
        return sum((R // seg + 1) + (R + 1) for R in range(T)) * nrhs
    n = 0
    for g in range((T + 3) // 4):
        last = min(4 * g + 3, T - 1)
        n += 4 * (last // seg + 1) + (last + 1)
    return n


def test_symv_workspace_covers_every_tile_pass_layout():
    """qc_symv_workspace_bytes (host arithmetic only) against the partial-sum
    blocks of both tile-pass layouts at every segment length they can use,
    down to a single tile row (where the grouped layout needs the most)."""
    from jax_qc import _native

    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libqcmd.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_native.LIB_PATH)
    lib.qc_symv_workspace_bytes.restype = ctypes.c_int64
    lib.qc_symv_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int]
    for nbf in list(range(1, 48)) + [120, 240, 480]:
        T = (nbf * (nbf + 1) // 2 + 63) // 64
        for nrhs in (1, 2, 3, 4):
            have = lib.qc_symv_workspace_bytes(nbf, nrhs)
            need = max(_tile_pass_blocks(T, s, nrhs, False) for s in range(1, 65))
            if nrhs == 1:
                need = max(need, max(_tile_pass_blocks(T, s, 1, True) for s in (4, 8, 12, 16)))
            assert have >= need * 64 * 8, (nbf, nrhs, have, need * 512)
