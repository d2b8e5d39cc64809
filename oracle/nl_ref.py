"""ORACLE -- test infrastructure only.  ctypes wrapper of nl_ref.c (the
float32 restatement of the neighbour search); builds it with gcc on first
use (``build()`` is also called by __graft_entry__.build())."""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "nl_ref.c")
OUT = os.path.join(HERE, "_build", "libnl_ref.so")
_lib = None


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", SRC, "-o",
           OUT + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(OUT) or os.path.getmtime(OUT) < os.path.getmtime(SRC):
            build()
        lib = ctypes.CDLL(OUT)
        lib.nlr_norm3_many.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        lib.nlr_build.restype = ctypes.c_int64
        lib.nlr_build.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_int64]
        _lib = lib
    return _lib


def norm3_f32(v: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.float32)
    out = np.empty(len(v), dtype=np.float32)
    _load().nlr_norm3_many(v.ctypes.data, len(v), out.ctypes.data)
    return out


def neighbor_csr(pos: np.ndarray, mol_ptr, cutoff: float, cells=None, pbc=None):
    """Receiver CSR (row_ptr int64 [N+1], col int32 [E], shift int32 [E,3])."""
    lib = _load()
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    mol_ptr = np.ascontiguousarray(mol_ptr, dtype=np.int32)
    n = len(pos)
    B = len(mol_ptr) - 1
    c = None if cells is None else np.ascontiguousarray(cells, dtype=np.float32).reshape(B, 3, 3)
    p = None if pbc is None else np.ascontiguousarray(pbc, dtype=np.uint8).reshape(B, 3)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    cap = 0
    E = lib.nlr_build(pos.ctypes.data, n, B, mol_ptr.ctypes.data,
                      None if c is None else c.ctypes.data, None if p is None else p.ctypes.data,
                      float(cutoff), row_ptr.ctypes.data, None, None, cap)
    col = np.empty(max(E, 1), dtype=np.int32)
    shift = np.empty((max(E, 1), 3), dtype=np.int32)
    lib.nlr_build(pos.ctypes.data, n, B, mol_ptr.ctypes.data,
                  None if c is None else c.ctypes.data, None if p is None else p.ctypes.data,
                  float(cutoff), row_ptr.ctypes.data, col.ctypes.data, shift.ctypes.data, E)
    return row_ptr, col[:E], shift[:E]
