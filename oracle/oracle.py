"""ctypes front end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  Never used by the product path.

Functions mirror the reference's host verbs on the flat arrays of a
``ChordalSymbolic`` (1-based Int64):
  * ``assemble``  — copyto!(F, A) (src/Multifrontal.jl/src/chordal_triangular.jl:852-893)
  * ``chol_impl`` — chol_impl!  (src/Multifrontal.jl/src/cholesky.jl:361-392)
  * ``solve``     — ldiv!(F, B) (src/Multifrontal.jl/src/divide.jl:191-206,1762-1764)
  * ``selinv``    — selinv!(F)  (src/Multifrontal.jl/src/selinv.jl:38-424)
  * ``complete``  — complete!(F) (src/Multifrontal.jl/src/complete.jl:1-451)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

i64p = ctypes.POINTER(ctypes.c_int64)
f64p = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_set_blas.argtypes = [ctypes.c_void_p] * 4
        L.or_set_blas.restype = None
        L.or_set_blas_syr.argtypes = [ctypes.c_void_p]
        L.or_set_blas_syr.restype = None
        L.or_chol_impl.argtypes = ([ctypes.c_int64] * 2 + [i64p] * 7 + [f64p] * 4 +
                                   [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, f64p, i64p, f64p])
        L.or_chol_impl.restype = ctypes.c_int64
        L.or_chol_range.argtypes = L.or_chol_impl.argtypes + [ctypes.c_int64, ctypes.c_int64]
        L.or_chol_range.restype = ctypes.c_int64
        for f in (L.or_div_fwd, L.or_div_bwd):
            f.argtypes = ([ctypes.c_int64] + [i64p] * 7 + [f64p] * 2 + [ctypes.c_int, f64p,
                          ctypes.c_int64, ctypes.c_int64, f64p, i64p, f64p])
            f.restype = None
        L.or_selinv.argtypes = [ctypes.c_int64] + [i64p] * 7 + [f64p] * 3 + [i64p, f64p, f64p]
        L.or_selinv.restype = ctypes.c_int64
        L.or_complete.argtypes = [ctypes.c_int64] + [i64p] * 7 + [f64p] * 3 + [i64p, f64p]
        L.or_complete.restype = ctypes.c_int64
        L.or_gmw81_beta.argtypes = [ctypes.c_int64] + [i64p] * 4 + [f64p] * 2
        L.or_gmw81_beta.restype = ctypes.c_double
        _lib = L
    return _lib


def _capsule_ptr(capsule):
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return get_ptr(capsule, get_name(capsule))


def use_host_blas(on=True):
    """Route the oracle's dense kernels (dpotrf, dtrsm, dsyrk, dgemm and the
    nn = 1 fronts' dsyr) to scipy's LAPACK/BLAS (OpenBLAS, LP64), as the
    reference does for Float64; ``on=False`` restores the C loops."""
    if not on:
        lib().or_set_blas(None, None, None, None)
        lib().or_set_blas_syr(None)
        return
    from scipy.linalg import cython_blas, cython_lapack
    b, l = cython_blas.__pyx_capi__, cython_lapack.__pyx_capi__
    lib().or_set_blas(_capsule_ptr(l["dpotrf"]), _capsule_ptr(b["dtrsm"]),
                      _capsule_ptr(b["dsyrk"]), _capsule_ptr(b["dgemm"]))
    lib().or_set_blas_syr(_capsule_ptr(b["dsyr"]))


def _p(a, t):
    return a.ctypes.data_as(t)


def _sym_args(S):
    return [_p(S.resptr, i64p), _p(S.sepptr, i64p), _p(S.reltgt, i64p), _p(S.chdptr, i64p),
            _p(S.chdtgt, i64p), _p(S.Dptr, i64p), _p(S.Lptr, i64p)]


def assemble(S, P, A):
    """Dval, Lval holding the lower triangle of P A P' (copyto!, uplo = :L)."""
    import scipy.sparse as sp
    A = sp.csc_matrix(A)
    n = A.shape[0]
    Dval = np.zeros(S.ndz)
    Lval = np.zeros(S.nlz)
    invp = P.invp - 1
    coo = A.tocoo()
    v = invp[coo.row]
    w = invp[coo.col]
    keep = v >= w
    v, w, x = v[keep], w[keep], coo.data[keep]
    j = S.idx[w] - 1
    r0 = S.resptr[j] - 1
    nn = S.resptr[j + 1] - S.resptr[j]
    inres = v < r0 + nn
    pd = S.Dptr[j] - 1 + (v - r0) + (w - r0) * nn
    np.add.at(Dval, pd[inres], x[inres])
    # rows in sep: locate with a search in septgt of the front
    nl = ~inres
    if nl.any():
        # septgt is sorted inside each front and fronts are stored in order, so
        # the key front*(n+1)+row is globally sorted.
        jj = j[nl]
        na_all = np.diff(S.sepptr)
        fr = np.repeat(np.arange(S.nfr, dtype=np.int64), na_all)
        keys = fr * (n + 1) + S.septgt
        q = jj * (n + 1) + (v[nl] + 1)
        g = np.searchsorted(keys, q)
        ok = (g < len(keys)) & (keys[np.minimum(g, len(keys) - 1)] == q)
        assert ok.all(), "entry outside the symbolic pattern"
        pos = g - (S.sepptr[jj] - 1)
        na = na_all[jj]
        pl = S.Lptr[jj] - 1 + pos + (w[nl] - r0[nl]) * na
        np.add.at(Lval, pl, x[nl])
    assert n == S.n
    return Dval, Lval


def gmw81_beta(S, Dval, Lval):
    """gmw81_beta (gmw81.jl:74-101) of the assembled factor storage."""
    Dval = np.ascontiguousarray(Dval, dtype=np.float64)
    Lval = np.ascontiguousarray(Lval, dtype=np.float64)
    return float(lib().or_gmw81_beta(S.nfr, _p(S.resptr, i64p), _p(S.sepptr, i64p), _p(S.Dptr, i64p),
                                     _p(S.Lptr, i64p), _p(Dval, f64p), _p(Lval, f64p)))


def chol_impl(S, Dval, Lval, diag=0, signs=None, reg=None):
    """In place on copies: returns (Dval, Lval, d, info); info as chol_impl!.
    reg: None, (delta, epsilon) for DynamicRegularization, or
    ("gmw81", beta, delta) for GMW81 (a negative beta takes gmw81_beta of the
    assembled matrix, a negative delta cbrt(eps))."""
    Dval = np.array(Dval, dtype=np.float64, copy=True)
    Lval = np.array(Lval, dtype=np.float64, copy=True)
    n = S.n
    d = np.ones(n) if diag == 0 else np.zeros(n)
    if signs is None:
        signs = np.ones(n) if diag == 0 else np.zeros(n)
    signs = np.ascontiguousarray(signs, dtype=np.float64)
    kind, delta, eps = 0, 0.0, 0.0
    if reg is not None and reg[0] == "gmw81":
        beta, dl = float(reg[1]), float(reg[2])
        kind, delta, eps = 2, (dl if dl >= 0 else float(np.cbrt(np.finfo(np.float64).eps))), beta
    elif reg is not None:
        kind, delta, eps = 1, float(reg[0]), float(reg[1])
    Mval = np.zeros(max(S.nMval, 1))
    Mptr = np.zeros(S.nMptr + 2, dtype=np.int64)
    Fval = np.zeros(max(S.nFval * S.nFval, 2))
    info = lib().or_chol_impl(n, S.nfr, *_sym_args(S), _p(Dval, f64p), _p(Lval, f64p),
                              _p(d, f64p), _p(signs, f64p), diag, kind, delta, eps,
                              _p(Mval, f64p), _p(Mptr, i64p), _p(Fval, f64p))
    return Dval, Lval, d, int(info)


def chol_subtree(S, Dval, Lval, root):
    """chol_impl!'s loop restricted to the subtree of front ``root`` (1-based):
    returns (seconds, reference flops of the subtree, info)."""
    import time
    nfr = S.nfr
    size = np.ones(nfr, dtype=np.int64)
    for j in range(nfr):
        p = S.pnt[j] - 1
        if p >= 0:
            size[p] += size[j]
    j1 = root
    j0 = root - size[root - 1]
    nn = S.nn()[j0:j1].astype(np.float64)
    na = S.na()[j0:j1].astype(np.float64)
    flops = float((nn ** 3 / 3 + na * nn ** 2 + na ** 2 * nn).sum())
    Dval = np.array(Dval, dtype=np.float64, copy=True)
    Lval = np.array(Lval, dtype=np.float64, copy=True)
    d = np.ones(S.n)
    signs = np.ones(S.n)
    Mval = np.zeros(max(S.nMval, 1))
    Mptr = np.zeros(S.nMptr + 2, dtype=np.int64)
    Fval = np.zeros(max(S.nFval * S.nFval, 2))
    t0 = time.perf_counter()
    info = lib().or_chol_range(S.n, S.nfr, *_sym_args(S), _p(Dval, f64p), _p(Lval, f64p),
                               _p(d, f64p), _p(signs, f64p), 0, 0, 0.0, 0.0,
                               _p(Mval, f64p), _p(Mptr, i64p), _p(Fval, f64p), j0, j1)
    return time.perf_counter() - t0, flops, int(info)


def div_natural(S, Dval, Lval, d, C, diag=0):
    """C <- L^-T [D^-1] L^-1 C for C in permuted order (n x nrhs, Fortran)."""
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    if C.ndim == 1:
        C = C.reshape(-1, 1, order="F")
    nrhs = C.shape[1]
    Mval = np.zeros(max(S.nNval * nrhs, 1))
    Mptr = np.zeros(S.nMptr + 2, dtype=np.int64)
    Fval = np.zeros(max(S.nFval * nrhs, 1))
    Dval = np.ascontiguousarray(Dval)
    Lval = np.ascontiguousarray(Lval)
    args = [S.nfr, *_sym_args(S), _p(Dval, f64p), _p(Lval, f64p), diag, _p(C, f64p),
            S.n, nrhs, _p(Mval, f64p), _p(Mptr, i64p), _p(Fval, f64p)]
    lib().or_div_fwd(*args)
    if diag == 1:
        C /= np.asarray(d)[:, None]
    lib().or_div_bwd(*args)
    return C


def solve(S, P, Dval, Lval, d, B, diag=0):
    """x = A \\ B through the factorization (ldiv!!: P, natural solve, P')."""
    B = np.asarray(B, dtype=np.float64)
    vec = B.ndim == 1
    C = P.apply(B)
    X = div_natural(S, Dval, Lval, d, C, diag)
    X = P.apply_inv(X)
    return X[:, 0] if vec else X


def selinv_stack_peak(S):
    """Peak of selinv_impl!'s update-matrix stack (elements): the root-first
    walk pops its own M and pushes one per child."""
    na = np.diff(S.sepptr)
    stack, top, peak = [], 0, 0
    for j in range(S.nfr - 1, -1, -1):
        if na[j] > 0:
            top -= stack.pop()
        for q in range(S.chdptr[j] - 1, S.chdptr[j + 1] - 1):
            i = S.chdtgt[q] - 1
            stack.append(int(na[i]) ** 2)
            top += stack[-1]
        peak = max(peak, top)
    return peak


def selinv(S, Dval, Lval):
    """selinv! of a Cholesky factor (uplo = :L) on copies: returns (Dval,
    Lval) holding (P A P')^-1 on the factor's pattern."""
    Dval = np.array(Dval, dtype=np.float64, copy=True)
    Lval = np.array(Lval, dtype=np.float64, copy=True)
    nn = np.diff(S.resptr)
    Mval = np.zeros(max(selinv_stack_peak(S), 1))
    Mptr = np.zeros(S.nfr + 2, dtype=np.int64)
    Fval = np.zeros(max(S.nFval * S.nFval, 2))
    Xval = np.zeros(max(int(nn.max()) ** 2, 1))
    lib().or_selinv(S.nfr, *_sym_args(S), _p(Dval, f64p), _p(Lval, f64p), _p(Mval, f64p),
                    _p(Mptr, i64p), _p(Fval, f64p), _p(Xval, f64p))
    return Dval, Lval


def complete(S, Dval, Lval):
    """complete! on copies: Dval/Lval hold a matrix on the factor's pattern
    (lower triangles); returns (Dval, Lval, info) with the Cholesky factor of
    the inverse of its maximum-determinant completion (strict upper triangles
    of the diagonal blocks set to 0)."""
    Dval = np.array(Dval, dtype=np.float64, copy=True)
    Lval = np.array(Lval, dtype=np.float64, copy=True)
    Mval = np.zeros(max(selinv_stack_peak(S), 1))
    Mptr = np.zeros(S.nfr + 2, dtype=np.int64)
    Fval = np.zeros(max(S.nFval * S.nFval, 2))
    info = lib().or_complete(S.nfr, *_sym_args(S), _p(Dval, f64p), _p(Lval, f64p), _p(Mval, f64p),
                             _p(Mptr, i64p), _p(Fval, f64p))
    return Dval, Lval, int(info)
