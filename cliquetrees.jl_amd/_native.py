"""Loading of the native libraries of the package.

``libmfsym.so`` (host C++) holds the symbolic analysis; ``libmfhip.so`` (HIP,
gfx950) holds the numeric engine.  Both are built in-tree by ``build.py``.
Nothing here falls back to Python: a missing library raises ``OSError`` at
the first call that needs it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SYM_PATH = os.path.join(HERE, "libmfsym.so")
HIP_PATH = os.path.join(HERE, "libmfhip.so")

_sym = None
_hip = None

i64p = ctypes.POINTER(ctypes.c_int64)
f64p = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p


def as_i64(a):
    """Contiguous int64 view of ``a`` plus its ctypes pointer."""
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(i64p)


def as_f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(f64p)


def sym():
    """The symbolic library (loaded once)."""
    global _sym
    if _sym is None:
        if not os.path.exists(SYM_PATH):
            raise OSError(f"{SYM_PATH} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(SYM_PATH)
        lib.mfs_symbolic.argtypes = [ctypes.c_int64, i64p, i64p, i64p, ctypes.c_int,
                                     ctypes.POINTER(vp)]
        lib.mfs_symbolic.restype = ctypes.c_int
        lib.mfs_sizes.argtypes = [vp, i64p]
        lib.mfs_sizes.restype = ctypes.c_int
        lib.mfs_export.argtypes = [vp] + [i64p] * 11
        lib.mfs_export.restype = ctypes.c_int
        lib.mfs_free.argtypes = [vp]
        lib.mfs_free.restype = None
        lib.mfs_flatindices.argtypes = [vp, ctypes.c_int64, i64p, i64p, i64p]
        lib.mfs_flatindices.restype = ctypes.c_int
        lib.mfs_mmd.argtypes = [ctypes.c_int64, i64p, i64p, ctypes.c_int64, i64p, i64p]
        lib.mfs_mmd.restype = ctypes.c_int
        lib.mfs_grid_nd.argtypes = [ctypes.c_int64] * 4 + [i64p]
        lib.mfs_grid_nd.restype = ctypes.c_int
        lib.mf_plan_check.argtypes = [ctypes.c_int64] * 2 + [i64p] * 8 + [i64p, ctypes.c_int64,
                                                                           ctypes.c_char_p,
                                                                           ctypes.c_int64]
        lib.mf_plan_check.restype = ctypes.c_int
        f64p = ctypes.POINTER(ctypes.c_double)
        lib.mf_selinv_check.argtypes = ([ctypes.c_int64] * 2 + [i64p] * 8 + [f64p] * 4 +
                                        [i64p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64])
        lib.mf_selinv_check.restype = ctypes.c_int
        lib.mf_cpl_solve_check.argtypes = [ctypes.c_int64, i64p, i64p, ctypes.c_uint64,
                                           ctypes.POINTER(ctypes.c_double), i64p, ctypes.c_int64,
                                           ctypes.c_char_p, ctypes.c_int64]
        lib.mf_cpl_solve_check.restype = ctypes.c_int
        _sym = lib
    return _sym


def hip():
    """The HIP numeric engine (loaded once).  Raises if it is not built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_PATH):
            raise OSError(f"{HIP_PATH} is not built; run __graft_entry__.build()")
        # PyTorch ships its own HIP runtime; load it first so that the engine
        # binds to that same runtime (one runtime per process, whichever of the
        # two the caller touches first)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(HIP_PATH)
        from . import _hipsig
        _hipsig.declare(lib)
        _hip = lib
    return _hip
