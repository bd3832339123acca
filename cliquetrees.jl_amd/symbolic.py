"""Host mirror of the reference's symbolic layer for the multifrontal path.

``symbolic(A, alg=..., snd=...)`` returns ``(Permutation, ChordalSymbolic)``
like ``symbolic(A; alg, snd)`` in
src/Multifrontal.jl/src/chordal_symbolic.jl:68-91; the arrays are computed by
libmfsym.so (C++ restatement of src/clique_trees.jl:106-234,
src/supernode_trees.jl:107-180 and chordal_symbolic.jl:302-378) and are kept
exactly as the reference stores them: Int64, 1-based.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _native

SND_KINDS = {"maximal": 0, "nodal": 1, "fundamental": 2}


@dataclass
class Permutation:
    """A permutation matrix P with ``(P*B)[k] = B[perm[k]]`` (1-based arrays).

    Mirrors ``Permutation`` (src/Multifrontal.jl/src/permutation.jl:12-33) and
    ``mul!(C, P, B)`` (src/Multifrontal.jl/src/multiply.jl:277-280).
    """

    perm: np.ndarray
    invp: np.ndarray

    @classmethod
    def from_perm(cls, perm):
        perm = np.asarray(perm, dtype=np.int64)
        invp = np.empty_like(perm)
        invp[perm - 1] = np.arange(1, len(perm) + 1, dtype=np.int64)
        return cls(perm, invp)

    def __len__(self):
        return len(self.perm)

    def apply(self, B):
        """P * B."""
        return np.asarray(B)[self.perm - 1]

    def apply_inv(self, C):
        """P' * C (= P \\ C)."""
        out = np.empty_like(np.asarray(C))
        out[self.perm - 1] = C
        return out


@dataclass
class ChordalSymbolic:
    """Flat arrays of ``ChordalSymbolic`` (chordal_symbolic.jl:15-28), 1-based Int64."""

    n: int
    nfr: int
    resptr: np.ndarray
    sepptr: np.ndarray
    septgt: np.ndarray
    reltgt: np.ndarray
    chdptr: np.ndarray
    chdtgt: np.ndarray
    pnt: np.ndarray
    idx: np.ndarray
    Dptr: np.ndarray
    Lptr: np.ndarray
    nMptr: int
    nMval: int
    nNval: int
    nFval: int
    _handle: object = None

    @property
    def ndz(self):
        return int(self.Dptr[-1] - 1)

    @property
    def nlz(self):
        return int(self.Lptr[-1] - 1)

    def nn(self):
        return np.diff(self.resptr)

    def na(self):
        return np.diff(self.sepptr)

    def res(self, j):
        """Columns of front j (1-based front index, 1-based columns)."""
        return np.arange(self.resptr[j - 1], self.resptr[j], dtype=np.int64)

    def sep(self, j):
        return self.septgt[self.sepptr[j - 1] - 1: self.sepptr[j] - 1]

    def rel(self, j):
        return self.reltgt[self.sepptr[j - 1] - 1: self.sepptr[j] - 1]

    def children(self, j):
        return self.chdtgt[self.chdptr[j - 1] - 1: self.chdptr[j] - 1]

    def factor_flops(self, ldlt=False):
        """Cholesky flops: sum nn^3/3 + na*nn^2 + na^2*nn (LDLt adds na*nn)."""
        nn = self.nn().astype(np.float64)
        na = self.na().astype(np.float64)
        f = nn ** 3 / 3.0 + na * nn ** 2 + na ** 2 * nn
        if ldlt:
            f = f + na * nn
        return float(f.sum())

    def solve_flops(self, nrhs=1):
        nn = self.nn().astype(np.float64)
        na = self.na().astype(np.float64)
        return float((2 * nn ** 2 + 4 * nn * na).sum()) * nrhs

    def __del__(self):
        h = self._handle
        if h is not None:
            try:
                _native.sym().mfs_free(h)
            except Exception:
                pass


def _csc(A):
    A = sp.csc_matrix(A)
    A.sort_indices()
    colptr = A.indptr.astype(np.int64) + 1
    rowval = A.indices.astype(np.int64) + 1
    return A, colptr, rowval


def grid_nd(nx, ny, nz=1, leaf=1):
    """Geometric nested-dissection elimination order of a box grid (1-based)."""
    n = nx * ny * nz
    order = np.empty(n, dtype=np.int64)
    rc = _native.sym().mfs_grid_nd(nx, ny, nz, leaf, order.ctypes.data_as(_native.i64p))
    if rc != 0:
        raise ValueError(f"mfs_grid_nd failed ({rc})")
    return order


def mmd(A, delta=0):
    """Multiple minimum degree order (``MMD()``): returns perm (1-based)."""
    A, colptr, rowval = _csc(A)
    n = A.shape[0]
    invp = np.empty(n, dtype=np.int64)
    perm = np.empty(n, dtype=np.int64)
    rc = _native.sym().mfs_mmd(n, colptr.ctypes.data_as(_native.i64p),
                               rowval.ctypes.data_as(_native.i64p), delta,
                               invp.ctypes.data_as(_native.i64p),
                               perm.ctypes.data_as(_native.i64p))
    if rc != 0:
        raise NotImplementedError(f"mfs_mmd failed ({rc})")
    return perm


def symbolic(A, alg=None, snd="maximal"):
    """``symbolic(A; alg, snd)``: returns ``(Permutation, ChordalSymbolic)``.

    ``alg`` is an elimination order (1-based vector, as accepted by the
    reference at src/elimination_algorithms.jl:776-779), the string ``"mmd"``,
    or None for the natural order.
    """
    A, colptr, rowval = _csc(A)
    n = A.shape[0]
    if A.shape[0] != A.shape[1]:
        raise ValueError("matrix must be square")
    if alg is None:
        order = np.arange(1, n + 1, dtype=np.int64)
    elif isinstance(alg, str):
        if alg.lower() != "mmd":
            raise ValueError(f"unknown ordering {alg!r}")
        order = mmd(A)
    else:
        order = np.ascontiguousarray(alg, dtype=np.int64)
        if order.shape != (n,):
            raise ValueError("permutation has the wrong length")
    lib = _native.sym()
    h = ctypes.c_void_p()
    rc = lib.mfs_symbolic(n, colptr.ctypes.data_as(_native.i64p),
                          rowval.ctypes.data_as(_native.i64p),
                          order.ctypes.data_as(_native.i64p), SND_KINDS[snd], ctypes.byref(h))
    if rc != 0:
        raise ValueError(f"mfs_symbolic failed ({rc})")
    sizes = np.zeros(9, dtype=np.int64)
    lib.mfs_sizes(h, sizes.ctypes.data_as(_native.i64p))
    n_, nfr, nsep = int(sizes[0]), int(sizes[1]), int(sizes[2])
    arr = {k: np.zeros(m, dtype=np.int64) for k, m in
           [("perm", n_), ("resptr", nfr + 1), ("sepptr", nfr + 1), ("septgt", nsep),
            ("reltgt", nsep), ("chdptr", nfr + 2), ("chdtgt", nfr), ("pnt", nfr),
            ("idx", n_), ("Dptr", nfr + 1), ("Lptr", nfr + 1)]}
    names = ["perm", "resptr", "sepptr", "septgt", "reltgt", "chdptr", "chdtgt", "pnt",
             "idx", "Dptr", "Lptr"]
    lib.mfs_export(h, *[arr[k].ctypes.data_as(_native.i64p) for k in names])
    S = ChordalSymbolic(n=n_, nfr=nfr, resptr=arr["resptr"], sepptr=arr["sepptr"],
                        septgt=arr["septgt"], reltgt=arr["reltgt"], chdptr=arr["chdptr"],
                        chdtgt=arr["chdtgt"], pnt=arr["pnt"], idx=arr["idx"],
                        Dptr=arr["Dptr"], Lptr=arr["Lptr"], nMptr=int(sizes[5]),
                        nMval=int(sizes[6]), nNval=int(sizes[7]), nFval=int(sizes[8]),
                        _handle=h)
    return Permutation.from_perm(arr["perm"]), S


def flatindices(S: ChordalSymbolic, A, uplo="L"):
    """``flatindices(S, A)`` (chordal_symbolic.jl:595-752): 1-based slot of each
    stored entry of A in [Dval; Lval], 0 if not in the stored triangle.  Indices
    follow the order of ``A.data`` of the CSC matrix with sorted indices.
    ``uplo="U"`` numbers the slots of the :U layout (the transposed blocks),
    where the stored triangle is the upper one."""
    A, colptr, rowval = _csc(A)
    if uplo == "U":
        # the upper entry (i, j) is stored where the :L layout stores (j, i),
        # at the transposed position inside its block
        coo = A.tocoo()
        T = sp.csc_matrix((np.arange(1, A.nnz + 1, dtype=np.float64), (coo.col, coo.row)),
                          shape=A.shape)
        T.sort_indices()
        flatT = flatindices(S, T, "L")
        src = T.data.astype(np.int64) - 1       # T's k-th entry is A's entry src[k]
        flat = np.zeros(A.nnz, dtype=np.int64)
        flat[src] = transpose_slots(S, flatT)
        return flat
    if uplo != "L":
        raise ValueError("uplo must be 'L' or 'U'")
    flat = np.zeros(len(rowval), dtype=np.int64)
    rc = _native.sym().mfs_flatindices(S._handle, A.shape[0],
                                       colptr.ctypes.data_as(_native.i64p),
                                       rowval.ctypes.data_as(_native.i64p),
                                       flat.ctypes.data_as(_native.i64p))
    if rc != 0:
        raise ValueError(f"mfs_flatindices failed ({rc})")
    return flat


def transpose_slots(S: ChordalSymbolic, flat):
    """Map 1-based slots of the :L layout to the slots of the same factor
    entries in the :U layout (0 stays 0); the map is its own inverse only
    on diagonal blocks, so it is used in the :L -> :U direction."""
    flat = np.asarray(flat, dtype=np.int64)
    out = np.zeros_like(flat)
    nn = np.diff(S.resptr)
    na = np.diff(S.sepptr)
    ndz = int(S.Dptr[-1] - 1)
    q = flat - 1
    inD = (flat > 0) & (q < ndz)
    if inD.any():
        j = np.searchsorted(S.Dptr - 1, q[inD], side="right") - 1
        e = q[inD] - (S.Dptr[j] - 1)
        r, c = e % nn[j], e // nn[j]
        out[inD] = S.Dptr[j] + c + r * nn[j]
    inL = (flat > 0) & (q >= ndz)
    if inL.any():
        ql = q[inL] - ndz
        j = np.searchsorted(S.Lptr - 1, ql, side="right") - 1
        e = ql - (S.Lptr[j] - 1)
        r, c = e % na[j], e // na[j]
        out[inL] = ndz + S.Lptr[j] + c + r * nn[j]
    return out
