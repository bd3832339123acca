"""Completion on the device (mf_complete_impl: each front's reordered matrix
factored as one root of a forest, the big ones on the tile-DAG, the result
read from the inverse of that factor) against the oracle's complete_impl!
restatement (pinned by complete!(selinv!(F)) = F and a dense closed form),
the same identity on the device at the 32^3 size, the failure info, both
layouts, and complete!(F) on a factor object."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from densehelp import dense_lower, permuted_dense, rel, to_layout, to_upper_layout

pytestmark = pytest.mark.gpu

RTOL = 1e-11


def _cases(ct):
    pr = ct.problems
    yield "lap2d_30_nd", pr.laplacian2d_dirichlet(30, 30), ct.grid_nd(30, 30, 1), "maximal"
    yield "lap3d_12_nd", pr.laplacian3d_7pt(12, 12, 12, seed=2), ct.grid_nd(12, 12, 12), "maximal"
    yield "st27_9_nd", pr.stencil3d_27pt(9, 9, 9, seed=3), ct.grid_nd(9, 9, 9), "maximal"
    yield "lap3d_10_mmd", pr.laplacian3d_7pt(10, 10, 10, seed=4), "mmd", "maximal"
    yield "lap3d_8_fund", pr.laplacian3d_7pt(8, 8, 8, seed=5), ct.grid_nd(8, 8, 8), "fundamental"
    # big fronts (nJ > 64): tile-DAG factors and k_selinv's inverses
    yield "lap3d_20_nd", pr.laplacian3d_7pt(20, 20, 20, seed=6), ct.grid_nd(20, 20, 20), "maximal"
    yield "st27_14_nd", pr.stencil3d_27pt(14, 14, 14, seed=7), ct.grid_nd(14, 14, 14), "maximal"


CASES = ["lap2d_30_nd", "lap3d_12_nd", "st27_9_nd", "lap3d_10_mmd", "lap3d_8_fund", "lap3d_20_nd",
         "st27_14_nd"]


@pytest.fixture(scope="module")
def cases(ct):
    return {name: rest for name, *rest in _cases(ct)}


def _selinv(ct, oracle, A, order, snd):
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D, L, d, info = oracle.chol_impl(S, Dv, Lv)
    SD, SL = oracle.selinv(S, D, L)
    return P, S, D, L, SD, SL


@pytest.mark.parametrize("name", CASES)
def test_complete_matches_oracle(ct, oracle, cases, name):
    A, order, snd = cases[name]
    print(f"complete case {name}", flush=True)
    P, S, D, L, SD, SL = _selinv(ct, oracle, A, order, snd)
    OD, OL, oinfo = oracle.complete(S, SD, SL)
    assert oinfo == 0
    CD, CL, info = ct.complete_impl(S, SD, SL)
    assert info == 0
    print(f"  vs oracle {rel(CD, OD):.2e} {rel(CL, OL):.2e}  vs factor {rel(CD, D):.2e} {rel(CL, L):.2e}")
    assert rel(CD, OD) <= RTOL and rel(CL, OL) <= RTOL
    # complete!(selinv!(F)) is F: the inverse of PAP' is the maximum-
    # determinant completion of its entries on the chordal pattern
    assert rel(CD, D) <= RTOL and rel(CL, L) <= RTOL
    # a second call on the same plan (kept work buffers), other values:
    # the completion of a scaled Sigma is the factor scaled by 1/sqrt(2)
    # (to the rounding of either computation)
    CD2, CL2, info2 = ct.complete_impl(S, 2.0 * SD, 2.0 * SL)
    assert info2 == 0 and rel(CD2, CD / np.sqrt(2.0)) <= RTOL and rel(CL2, CL / np.sqrt(2.0)) <= RTOL


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_complete_impl_layouts(ct, oracle, cases, uplo):
    """complete_impl! on host arrays in either layout (the strict other
    triangle of each diagonal block is 0, as in the oracle)."""
    A, order, snd = cases["lap3d_20_nd"]
    print(f"complete uplo={uplo}", flush=True)
    P, S, D, L, SD, SL = _selinv(ct, oracle, A, order, snd)
    OD, OL, _ = oracle.complete(S, SD, SL)
    if uplo == "U":
        SD, SL = to_upper_layout(S, SD, SL)
        OD, OL = to_upper_layout(S, OD, OL)
    CD, CL, info = ct.complete_impl(S, SD, SL, uplo=uplo)
    assert info == 0 and rel(CD, OD) <= RTOL and rel(CL, OL) <= RTOL


def test_complete_of_a_dense_pattern(ct):
    """One front (a dense 6 x 6 pattern): the completion is the matrix itself,
    the result the Cholesky factor of its inverse."""
    rng = np.random.default_rng(3)
    G = rng.standard_normal((6, 6))
    Sig = G @ G.T + 6 * np.eye(6)
    P, S = ct.symbolic(sp.csc_matrix(np.ones((6, 6))))
    Sp = permuted_dense(Sig, P)
    Dv, Lv = to_layout(S, np.tril(Sp))
    CD, CL, info = ct.complete_impl(S, Dv, Lv)
    assert info == 0
    assert rel(dense_lower(S, CD, CL), np.linalg.cholesky(np.linalg.inv(Sp))) <= 1e-13


@pytest.mark.parametrize("name", ["lap2d_30_nd", "lap3d_20_nd"])
def test_complete_failure_info(ct, oracle, cases, name):
    """A negated selected inverse: the first potrf! in the reference's order
    of fronts fails; info is the oracle's (the failing local column) and the
    arrays come back as given."""
    A, order, snd = cases[name]
    print(f"complete failure {name}", flush=True)
    P, S, D, L, SD, SL = _selinv(ct, oracle, A, order, snd)
    _, _, oinfo = oracle.complete(S, -SD, -SL)
    CD, CL, info = ct.complete_impl(S, -SD, -SL)
    assert oinfo > 0 and info == oinfo
    assert np.array_equal(CD, -SD) and np.array_equal(CL, -SL)


def test_complete_factor_object(ct, oracle, cases):
    """complete!(selinv!(cholesky!(F))) on a factor object: F is its own
    factor again, and solves with it."""
    A, order, snd = cases["st27_14_nd"]
    print("complete factor object", flush=True)
    F = ct.ChordalCholesky(A, alg=order, snd=snd)
    ct.cholesky_(F)
    D0, L0 = F.Dval, F.Lval
    ct.selinv_(F)
    SD, SL = F.Dval, F.Lval
    ct.complete_(F)
    assert F.info == 0
    assert rel(F.Dval, D0) <= RTOL and rel(F.Lval, L0) <= RTOL
    b = np.random.default_rng(4).standard_normal(A.shape[0])
    x = ct.solve(F, b)
    assert np.abs(A @ x - b).max() <= 1e-9 * np.abs(b).max()
    # failure: check=True raises PosDefException with the info
    G = ct.ChordalCholesky(A, F.P, F.S)
    G.set_values(-SD, -SL)
    with pytest.raises(ct.PosDefException):
        ct.complete_(G)
    assert G.info > 0


def test_complete_factor_object_upper(ct, cases):
    """complete!(selinv!(cholesky!(F))) on a ChordalCholesky{:U} factor
    object: its Dval/Lval hold the transposed blocks throughout, and F is
    its own factor again."""
    A, order, snd = cases["lap3d_20_nd"]
    print("complete factor object uplo=U", flush=True)
    F = ct.ChordalCholesky(A, alg=order, snd=snd, uplo="U")
    ct.cholesky_(F)
    D0, L0 = F.Dval, F.Lval
    ct.selinv_(F)
    ct.complete_(F)
    assert F.info == 0
    assert rel(F.Dval, D0) <= RTOL and rel(F.Lval, L0) <= RTOL
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    assert np.abs(A @ ct.solve(F, b) - b).max() <= 1e-9 * np.abs(b).max()


def test_complete_in_chunks(ct, oracle, cases):
    """The forest cut into about ten chunks (a small memory budget through
    MF_COMPLETE_CHUNK_ELEMS, read when a plan's completion is first built):
    the same values as one chunk, and a failure in a leaf front (its
    diagonal block negated, so only its own Schur complement is indefinite)
    found in its chunk with the oracle's info."""
    A, order, snd = cases["lap3d_20_nd"]
    print("complete in chunks", flush=True)
    P, S, D, L, SD, SL = _selinv(ct, oracle, A, order, snd)
    nj = np.diff(S.resptr) + np.diff(S.sepptr)
    CD1, CL1, _ = ct.complete_impl(S, SD, SL)
    P2, S2 = ct.symbolic(A, alg=order, snd=snd)
    os.environ["MF_COMPLETE_CHUNK_ELEMS"] = str(int((nj * nj).sum()) // 10)
    try:
        CD, CL, info = ct.complete_impl(S2, SD, SL)
        assert info == 0
        assert np.array_equal(CD, CD1) and np.array_equal(CL, CL1)
        # a leaf: a front without children
        leaves = np.flatnonzero(np.diff(S.chdptr[: S.nfr + 1]) == 0)
        j = int(leaves[len(leaves) // 2])
        nn = int(S.resptr[j + 1] - S.resptr[j])
        BD = SD.copy()
        BD[S.Dptr[j] - 1: S.Dptr[j] - 1 + nn * nn] *= -1.0
        _, _, oinfo = oracle.complete(S, BD, SL)
        XD, XL, info = ct.complete_impl(S2, BD, SL)
        assert oinfo > 0 and info == oinfo
        assert np.array_equal(XD, BD) and np.array_equal(XL, SL)
    finally:
        del os.environ["MF_COMPLETE_CHUNK_ELEMS"]


def test_complete_lap7_32_round_trip(ct):
    """32^3 7-point Laplacian, all on the device: selinv! then complete!
    gives the factor back (a property that needs no oracle), at a size with
    several levels of big fronts."""
    A = ct.problems.laplacian3d_7pt(32, 32, 32, seed=8)
    print("complete lap7_32", flush=True)
    F = ct.ChordalCholesky(A, alg=ct.grid_nd(32, 32, 32))
    ct.cholesky_(F)
    D0, L0 = F.Dval, F.Lval
    ct.selinv_(F)
    ct.complete_(F)
    assert rel(F.Dval, D0) <= 1e-10 and rel(F.Lval, L0) <= 1e-10
