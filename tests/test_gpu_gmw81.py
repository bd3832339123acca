"""GMW81 on the device (MF_REG_GMW81: k_small / wave fronts with the column
maximum by wavefront reduction, the tile-DAG's diagonal tiles by potrf_gmw
with the rows below the tile, the DAG in creation order) against the
oracle's GMW81 restatement (pinned by the docstring, a dense restatement and
the method's invariants), the docstring's printed values, and properties of
the regularized factor."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from densehelp import dense_lower, permuted_dense, rel
from test_oracle_gmw81 import GMW_A, GMW_B, GMW_D, _shifted

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def _cases(ct):
    pr = ct.problems
    yield "lap2d_20_mmd", _shifted(pr.laplacian2d_dirichlet(20, 20), 0.3), "mmd", "maximal"
    yield "lap2d_16_fund", _shifted(pr.laplacian2d_dirichlet(16, 16), 0.1), "mmd", "fundamental"
    yield "lap3d_10_nd", _shifted(pr.laplacian3d_7pt(10, 10, 10, seed=2), 0.2), ct.grid_nd(10, 10, 10), "maximal"
    # big fronts (nJ > 64): the tile-DAG with several block columns and rows below
    yield "lap3d_16_nd", _shifted(pr.laplacian3d_7pt(16, 16, 16, seed=5), 0.05), ct.grid_nd(16, 16, 16), "maximal"
    yield "st27_12_nd", _shifted(pr.stencil3d_27pt(12, 12, 12, seed=3), 0.02), ct.grid_nd(12, 12, 12), "maximal"
    yield "lap3d_20x18x9_nd", _shifted(pr.laplacian3d_7pt(20, 18, 9, seed=6), 0.1), ct.grid_nd(20, 18, 9), "maximal"


CASES = ["lap2d_20_mmd", "lap2d_16_fund", "lap3d_10_nd", "lap3d_16_nd", "st27_12_nd", "lap3d_20x18x9_nd"]


@pytest.fixture(scope="module")
def cases(ct):
    return {name: rest for name, *rest in _cases(ct)}


def _oracle(oracle, S, P, A, diag, signs=None, beta=-1.0, delta=-1.0):
    Dv, Lv = oracle.assemble(S, P, A)
    sp_ = None if signs is None else np.ascontiguousarray(np.asarray(signs, float)[P.perm - 1])
    return oracle.chol_impl(S, Dv, Lv, diag, signs=sp_, reg=("gmw81", beta, delta))


@pytest.mark.parametrize("order", [[1, 2, 3], [3, 1, 2]])
def test_gmw81_docstring_on_device(ct, oracle, order):
    """gmw81.jl:8-30: Matrix(F) as printed; diag(F) of the reference test."""
    print(f"gmw81 docstring order={order}", flush=True)
    A = sp.csc_matrix(GMW_A)
    P, S = ct.symbolic(A, alg=np.array(order))
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F, reg=ct.GMW81())
    assert ct.issuccess(F)
    D0, L0, _, _ = _oracle(oracle, S, P, A, 0)
    assert rel(F.Dval, D0) <= RTOL64 and rel(F.Lval, L0) <= RTOL64
    G = ct.ChordalLDLt(A, P, S)
    ct.ldlt_(G, signs=np.ones(3), reg=ct.GMW81())
    _, _, d1, _ = _oracle(oracle, S, P, A, 1, np.ones(3))
    assert rel(G.d, d1) <= RTOL64
    if order == [1, 2, 3]:
        Lf = dense_lower(S, F.Dval, F.Lval)
        assert np.allclose(Lf @ Lf.T, GMW_B, rtol=0, atol=6e-6)
        assert np.allclose(G.d, GMW_D, rtol=1e-3)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("diag", [0, 1])
def test_gmw81_matches_oracle(ct, oracle, cases, name, diag):
    A, order, snd = cases[name]
    print(f"gmw81 case {name} diag={diag}", flush=True)
    n = A.shape[0]
    P, S = ct.symbolic(A, alg=order, snd=snd)
    signs = np.ones(n) if diag else None
    D0, L0, d0, info0 = _oracle(oracle, S, P, A, diag, signs)
    assert info0 == 0
    if diag:
        F = ct.ChordalLDLt(A, P, S)
        ct.ldlt_(F, signs=signs, reg=ct.GMW81())
        assert rel(F.d, d0) <= RTOL64
    else:
        F = ct.ChordalCholesky(A, P, S)
        ct.cholesky_(F, reg=ct.GMW81())
    assert ct.issuccess(F)
    assert rel(F.Lval, L0) <= RTOL64 and rel(F.Dval, D0) <= RTOL64
    # the solve runs on the regularized factor: it solves (PAP' + E) y = Pb
    # with E diagonal, backward stably (normwise backward error in the
    # infinity norm; the regularized matrices here are far from well scaled)
    b = np.random.default_rng(2).standard_normal(n)
    x = ct.solve(F, b)
    if n <= 3000:
        Lf = dense_lower(S, D0, L0, unit=diag == 1)
        M = (Lf * d0) @ Lf.T if diag else Lf @ Lf.T
        xp, bp = x[P.perm - 1], b[P.perm - 1]
        r = np.abs(M @ xp - bp).max()
        assert r <= 1e-13 * (np.abs(M).sum(axis=1).max() * np.abs(xp).max() + np.abs(bp).max())


def test_gmw81_ldlt_kkt_signs(ct, oracle):
    K, nh, m = ct.problems.kkt_matrix(ct.problems.near_cubic_dims(1500, 3), seed=7)
    print("gmw81 kkt", flush=True)
    n = K.shape[0]
    signs = np.concatenate([np.ones(nh), -np.ones(n - nh)])
    P, S = ct.symbolic(K, alg="mmd")
    D0, L0, d0, info0 = _oracle(oracle, S, P, K, 1, signs)
    F = ct.ChordalLDLt(K, P, S)
    ct.ldlt_(F, signs=signs, reg=ct.GMW81())
    assert rel(F.d, d0) <= RTOL64 and rel(F.Lval, L0) <= RTOL64 and rel(F.Dval, D0) <= RTOL64
    assert np.all(np.sign(F.d[P.invp - 1]) == signs)


def test_gmw81_parameters(ct, oracle, cases):
    """Given beta and delta (no device reduction for beta)."""
    A, order, snd = cases["lap3d_16_nd"]
    print("gmw81 parameters", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    D0, L0, _, _ = _oracle(oracle, S, P, A, 0, beta=3.0, delta=0.25)
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F, reg=ct.GMW81(beta=3.0, delta=0.25))
    assert rel(F.Lval, L0) <= RTOL64 and rel(F.Dval, D0) <= RTOL64


def test_gmw81_fp32(ct, oracle, cases):
    A, order, snd = cases["lap3d_16_nd"]
    print("gmw81 fp32", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    F = ct.ChordalCholesky(A, P, S, precision=ct.MF_FP32)
    ct.cholesky_(F, reg=ct.GMW81())
    D0, L0, _, _ = _oracle(oracle, S, P, A, 0)
    # FP32 (delta = cbrt(eps(Float32)), beta rounded to Float32): within
    # single-precision rounding of the FP64 oracle on a mildly shifted matrix
    assert rel(F.Lval, L0) <= 1e-3 and rel(F.Dval, D0) <= 1e-3


def test_gmw81_repeat_bitwise(ct, cases):
    """Two factorizations of the same values give the same bits (the DAG's
    order of work does not change the arithmetic)."""
    A, order, snd = cases["st27_12_nd"]
    print("gmw81 repeat", flush=True)
    F = ct.ChordalCholesky(A, alg=order, snd=snd)
    ct.cholesky_(F, reg=ct.GMW81())
    D1, L1 = F.Dval.copy(), F.Lval.copy()
    F.copyto_(A)
    ct.cholesky_(F, reg=ct.GMW81())
    assert np.array_equal(F.Dval, D1) and np.array_equal(F.Lval, L1)


def test_gmw81_needs_signs(ct, cases):
    A, order, snd = cases["lap2d_20_mmd"]
    F = ct.ChordalLDLt(A, alg=order, snd=snd)
    with pytest.raises(ValueError):
        ct.ldlt_(F, reg=ct.GMW81())
    from cliquetrees_amd import _native
    lib = _native.hip()
    info = ctypes.c_int64(0)
    rc = lib.mf_factorize(F._f.handle, ct.MF_LDLT, None, 2, -1.0, -1.0, ctypes.byref(info))
    assert rc == -1 and b"signs" in lib.mf_last_error()


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_gmw81_dropin_chol_impl(ct, oracle, cases, uplo):
    """chol_impl! with GMW81 on host arrays (mf_chol_impl) in either layout,
    Cholesky and LDLt with signs (permuted order, as chol_impl! receives them)."""
    from densehelp import to_upper_layout
    A, order, snd = cases["lap3d_16_nd"]
    print(f"gmw81 drop-in uplo={uplo}", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    for diag in (0, 1):
        sp_ = np.ones(S.n) if diag else None
        D0, L0, d0, _ = oracle.chol_impl(S, Dv, Lv, diag, signs=sp_, reg=("gmw81", -1.0, -1.0))
        Din, Lin = (Dv, Lv) if uplo == "L" else to_upper_layout(S, Dv, Lv)
        D1, L1, d1, info = ct.chol_impl(S, Din, Lin, diag=diag, signs=sp_, reg=ct.GMW81(), uplo=uplo)
        if uplo == "U":
            D0, L0 = to_upper_layout(S, D0, L0)
        assert info == 0 and rel(D1, D0) <= RTOL64 and rel(L1, L0) <= RTOL64
        if diag:
            assert rel(d1, d0) <= RTOL64


def test_gmw81_perturbation_at_size(ct):
    """32 x 32 x 30 7-point Laplacian shifted into indefiniteness, GMW81
    Cholesky on the device: L L' - PAP' is diagonal (to rounding) with
    nonnegative entries, and some pivots were raised (a property of the
    method that needs no oracle, at a size with several levels of big
    fronts)."""
    import scipy.sparse as sp
    from densehelp import sparse_lower
    A = _shifted(ct.problems.laplacian3d_7pt(32, 32, 30, seed=9), 0.02)
    print("gmw81 perturbation 32x32x30", flush=True)
    P, S = ct.symbolic(A, alg=ct.grid_nd(32, 32, 30))
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F, reg=ct.GMW81())
    L = sparse_lower(S, F.Dval, F.Lval)
    p = P.perm - 1
    M = sp.csc_matrix(A)[p][:, p]
    E = (L @ L.T - M).tocoo()
    off = np.abs(E.data[E.row != E.col]).max()
    # rounding of the products L L' (backward error of a Cholesky
    # factorization: |E| <= c n eps |L| |L'|), entries of the regularized
    # factor being larger than those of A
    absL = abs(L)
    scale = (absL @ absL.T).max()
    d = E.diagonal()
    assert off <= 64 * np.finfo(float).eps * scale
    assert d.min() >= -64 * np.finfo(float).eps * scale and d.max() > 1e-3
