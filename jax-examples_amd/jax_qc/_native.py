"""ctypes binding of libqcmd.so (the HIP C-ABI declared in include/qcmd.h).

There is no fallback: if the library is missing or no GPU is visible, every
device entry point raises.  Array arguments are torch tensors on the GPU;
the current torch stream is passed to every launch.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libqcmd.so")

_lib = None
_lock = threading.Lock()
_inited = set()

c_int, c_i64, c_double, c_void_p = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class QcEriArgs(ctypes.Structure):
    """Mirror of qc_eri_args (include/qcmd.h)."""

    _fields_ = [
        ("pairs64", c_void_p), ("pairs32", c_void_p), ("sp_prim_off", c_void_p),
        ("sp_nprim", c_void_p), ("sp_ao", c_void_p), ("bra_list", c_void_p),
        ("ket_list", c_void_p), ("offsets", c_void_p), ("kstart", c_void_p),
        ("nbra", c_int), ("nquartets", c_i64), ("mode", c_int), ("nbf", c_i64),
        ("out", c_void_p), ("schwarz", c_void_p), ("boys64", c_void_p), ("boys32", c_void_p),
        ("variant", c_int), ("sp_base", c_void_p), ("block", c_int),
    ]


class QcGrad2eArgs(ctypes.Structure):
    """Mirror of qc_grad2e_args (include/qcmd.h)."""

    _fields_ = [
        ("pairs64", c_void_p), ("sp_prim_off", c_void_p), ("sp_nprim", c_void_p),
        ("sp_aoc", c_void_p), ("sp_atom", c_void_p), ("sp_diag", c_void_p),
        ("bra_list", c_void_p), ("ket_list", c_void_p), ("offsets", c_void_p),
        ("kstart", c_void_p), ("nbra", c_int), ("nquartets", c_i64), ("ncart", c_i64),
        ("dj", c_void_p), ("dk0", c_void_p), ("dk1", c_void_p), ("nk", c_int),
        ("cx", c_double), ("grad", c_void_p), ("variant", c_int), ("schwarz", c_void_p),
        ("tau_density", c_double), ("pairs32", c_void_p), ("fp32", c_int), ("natom", c_int),
    ]


_SIGS = {
    "qc_init": [c_int],
    "qc_boys": [c_int, c_void_p, c_void_p, c_i64, c_void_p],
    "qc_boys_f32": [c_int, c_void_p, c_void_p, c_i64, c_void_p],
    "qc_eri": [c_int, c_int, c_int, ctypes.POINTER(QcEriArgs), c_void_p],
    "qc_one_electron": [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                        c_int, c_void_p, c_void_p, c_i64, c_void_p, c_void_p],
    "qc_sm_assemble": [c_int, c_void_p, c_void_p, c_i64, c_void_p],
    "qc_sm_assemble_maxabs": [c_int, c_void_p, c_void_p, c_i64, c_void_p, c_void_p],
    "qc_sm_p_from_jk": [c_void_p, c_void_p, c_i64, c_void_p, c_void_p],
    "qc_symv": [c_void_p, c_i64, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "qc_symv_tiles": [c_void_p, c_i64, c_int, c_void_p, c_void_p, c_void_p],
    "qc_symv_reduce": [c_i64, c_int, c_void_p, c_void_p, c_void_p],
    "qc_sm_tile_maxabs": [c_void_p, c_i64, c_void_p, c_void_p],
    "qc_sm_pack": [c_void_p, c_i64, c_void_p, c_void_p, c_void_p, c_void_p],
    "qc_symv_mixed_tiles": [c_void_p, c_void_p, c_void_p, c_i64, c_int, c_void_p, c_void_p,
                            c_void_p],
    "qc_symv_config": [c_int, c_int],
    "qc_symv_split": [c_int, c_int],
    "qc_symv_noalloc": [c_void_p, c_i64, c_int, c_void_p, c_void_p, c_void_p],
    "qc_block_maxabs": [c_void_p, c_i64, c_int, c_void_p, c_void_p],
    "qc_symv_screened_tiles": [c_void_p, c_void_p, c_void_p, c_void_p, c_i64, c_int, c_void_p,
                               c_void_p, c_double, c_void_p, c_void_p],
    "qc_symv_screened_reduce": [c_i64, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_double, c_void_p],
    "qc_pack_density": [c_void_p, c_i64, c_int, c_void_p, c_void_p],
    "qc_unpack_fock": [c_void_p, c_void_p, c_double, c_i64, c_int, c_void_p, c_void_p],
    "qc_dense_jk": [c_void_p, c_void_p, c_i64, c_void_p, c_void_p, c_void_p],
    "qc_grad2e": [c_int, c_int, ctypes.POINTER(QcGrad2eArgs), c_void_p],
    "qc_grad2e_prescreen": [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                            c_void_p, c_void_p, c_void_p, ctypes.c_double, c_void_p,
                            ctypes.c_double, c_void_p, c_void_p, c_void_p, c_void_p],
    "qc_grad1e": [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                  c_int, c_int, c_void_p, c_void_p, c_i64, c_void_p, c_void_p, c_void_p,
                  c_void_p],
    "qc_tiles_maxabs": [c_void_p, c_i64, c_void_p, c_void_p],
    "qc_tc2_partials": [c_void_p, c_i64, c_int, c_void_p, c_void_p],
    "qc_tc2_update": [c_void_p, c_void_p, c_void_p, c_void_p, c_i64, c_int, c_void_p, c_void_p,
                      c_double, c_void_p],
    "qc_tc2_tile_partials": [c_void_p, c_i64, c_int, c_void_p, c_void_p],
    "qc_tc2_step": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_i64, c_int, c_void_p,
                    c_double, c_void_p],
    "qc_one_electron_batch": [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_int, c_void_p, c_void_p, c_void_p, c_i64, c_void_p, c_void_p,
                              c_void_p],
    "qc_tiles_pack": [c_void_p, c_i64, c_void_p, c_void_p, c_void_p, c_void_p],
    "qc_batch_symv_tiles": [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_i64, c_i64, c_i64, c_int, c_int, c_void_p, c_void_p,
                            c_void_p],
    "qc_batch_symv_reduce": [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_i64, c_i64, c_i64,
                             c_int, c_int, c_void_p, c_void_p, c_void_p],
    "qc_batch_pack_density": [c_void_p, c_i64, c_int, c_void_p, c_void_p, c_int, c_void_p, c_i64,
                              c_void_p],
    "qc_batch_unpack_fock": [c_void_p, c_void_p, c_double, c_i64, c_int, c_void_p, c_void_p,
                             c_int, c_i64, c_void_p, c_void_p],
}
_OPTIONAL_SIGS = {}


def register_optional(name, argtypes, restype=c_int):
    _OPTIONAL_SIGS[name] = (argtypes, restype)


def lib():
    """Load libqcmd.so (raises if it was not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                    "g.build()'` (hipcc, gfx950). There is no CPU fallback.")
            L = ctypes.CDLL(LIB_PATH)
            for name, argt in _SIGS.items():
                fn = getattr(L, name)
                fn.argtypes = argt
                fn.restype = c_int
            L.qc_last_error.restype = ctypes.c_char_p
            L.qc_last_error.argtypes = []
            L.qc_symv_workspace_bytes.restype = c_i64
            L.qc_symv_workspace_bytes.argtypes = [c_i64, c_int]
            L.qc_tc2_partials_len.restype = c_i64
            L.qc_tc2_partials_len.argtypes = [c_int]
            L.qc_tc2_tile_partials_len.restype = c_i64
            L.qc_tc2_tile_partials_len.argtypes = [c_i64, c_int]
            L.qc_batch_items.restype = c_i64
            L.qc_batch_items.argtypes = [ctypes.POINTER(c_i64), c_int, c_int,
                                         ctypes.POINTER(c_i64)]
            for name, (argt, rt) in _OPTIONAL_SIGS.items():
                fn = getattr(L, name)
                fn.argtypes = argt
                fn.restype = rt
            _lib = L
    return _lib


def exported_symbols():
    return (list(_SIGS) + ["qc_last_error", "qc_symv_workspace_bytes", "qc_batch_items",
                           "qc_tc2_partials_len", "qc_tc2_tile_partials_len"]
            + list(_OPTIONAL_SIGS))


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("jax_qc device path needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def ensure_init(dev: torch.device | None = None):
    dev = dev or device()
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    if idx not in _inited:
        with torch.cuda.device(idx):
            call("qc_init", idx)
        _inited.add(idx)


def stream_ptr():
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    if t is None:
        return c_void_p(0)
    if isinstance(t, torch.Tensor):
        if not t.is_cuda:
            raise ValueError("device tensor expected")
        if not t.is_contiguous():
            raise ValueError("contiguous tensor expected")
        return c_void_p(t.data_ptr())
    return t


def call(name, *args):
    fn = getattr(lib(), name)
    rc = fn(*args)
    if rc != 0:
        msg = lib().qc_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed: {msg}")
    return rc
