"""Batches of independent sparse systems of uneven size.

The instances are factored as one forest: the block-diagonal matrix with the
instances' own elimination orders (offset into the block) has one elimination
tree per instance, so one plan covers the batch.  The persistent small-front
kernel then walks the levels of all trees in one launch, and the tile-DAG
kernel takes the big fronts of every instance in one launch; nothing in the
schedule couples two instances.  Each instance's factor is the factor it would
have alone (same order, same supernodes, same arithmetic).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .factorization import (MF_CHOLESKY, MF_FP64, MF_LDLT, ChordalCholesky, ChordalLDLt,
                            PosDefException, SingularException, cholesky_, ldlt_, refine, solve)
from .symbolic import symbolic


class BatchFailure(Exception):
    """A numeric failure inside a batch: ``instance`` is the failing system
    (0-based), ``info`` its failing vertex (1-based, the system's own
    numbering: what F.info holds for the system alone, cholesky.jl:250-254)."""

    def __init__(self, instance, info, cause):
        super().__init__(f"instance {instance}: {cause!r} at column {info}")
        self.instance, self.info, self.cause = instance, info, cause


class BatchedFactorization:
    """Cholesky (or LDLt) of many independent systems at once.

    mats:    list of square sparse matrices (SPD, or quasi-definite for LDLt)
    orders:  per-instance elimination orders (1-based), or None (natural)
    """

    def __init__(self, mats, orders=None, precision=MF_FP64, diag=MF_CHOLESKY):
        if not mats:
            raise ValueError("empty batch")
        self.sizes = np.array([m.shape[0] for m in mats], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        A = sp.block_diag([sp.csc_matrix(m, dtype=np.float64) for m in mats], format="csc")
        A.sort_indices()
        if orders is None:
            order = None
        else:
            if len(orders) != len(mats):
                raise ValueError("one order per instance")
            parts = []
            for o, off, k in zip(orders, self.offsets[:-1], self.sizes):
                o = np.asarray(o, dtype=np.int64)
                if o.shape != (k,):
                    raise ValueError("order has the wrong length")
                parts.append(o + off)
            order = np.concatenate(parts)
        self.diag = diag
        self.P, self.S = symbolic(A, alg=order)
        cls = ChordalCholesky if diag == MF_CHOLESKY else ChordalLDLt
        self.F = cls(A, self.P, self.S, precision=precision)
        self.A = A

    @property
    def count(self):
        return len(self.sizes)

    @property
    def n(self):
        return int(self.offsets[-1])

    def update_values(self, mats):
        """New values on the same patterns (device-side setflatindex!)."""
        A = sp.block_diag([sp.csc_matrix(m, dtype=np.float64) for m in mats], format="csc")
        A.sort_indices()
        self.F.copyto_(A)
        self.A = A
        return self

    def factorize(self, signs=None, reg=None):
        """cholesky!/ldlt! of every instance; raises BatchFailure naming the
        first failing instance (lowest failing global column)."""
        try:
            if self.diag == MF_CHOLESKY:
                cholesky_(self.F, reg=reg)
            else:
                ldlt_(self.F, signs=signs, reg=reg)
        except (PosDefException, SingularException) as e:
            g = int(self.F.perm[e.info - 1])
            k = int(np.searchsorted(self.offsets, g - 1, side="right") - 1)
            raise BatchFailure(k, g - int(self.offsets[k]), e) from e
        return self

    def _stack(self, bs):
        if isinstance(bs, np.ndarray):
            B = np.asarray(bs, dtype=np.float64)
            if B.shape[0] != self.n:
                raise ValueError("dimension mismatch")
            return B, False
        if len(bs) != self.count:
            raise ValueError("one right-hand side per instance")
        return np.concatenate([np.asarray(b, dtype=np.float64) for b in bs], axis=0), True

    def _split(self, X, as_list):
        if not as_list:
            return X
        return [X[self.offsets[k]:self.offsets[k + 1]] for k in range(self.count)]

    def solve(self, bs):
        """Per-instance solutions (list in, list out; or the stacked array)."""
        B, as_list = self._stack(bs)
        return self._split(solve(self.F, B), as_list)

    def refine(self, bs, tol=1e-10, maxit=20):
        """Iterative refinement until every system meets tol (FP64 residual)."""
        B, as_list = self._stack(bs)
        return self._split(refine(self.F, B, tol=tol, maxit=maxit, segments=self.offsets), as_list)


def batched_cholesky(mats, orders=None, precision=MF_FP64):
    return BatchedFactorization(mats, orders, precision, MF_CHOLESKY).factorize()


def batched_ldlt(mats, orders=None, precision=MF_FP64, signs=None, reg=None):
    return BatchedFactorization(mats, orders, precision, MF_LDLT).factorize(signs=signs, reg=reg)


__all__ = ["BatchedFactorization", "BatchFailure", "batched_cholesky", "batched_ldlt", "MF_LDLT",
           "MF_FP64"]
