"""Selected inversion on the device (mf_selinv, k_selinv) against the oracle's
selinv_impl! restatement (pinned by the selinv! docstring and dense inverses),
the docstring's printed inverse itself, dense inverses, the drop-in
mf_selinv_impl in both layouts, FP32 fronts, and at the 64^3 size a property
that needs no oracle: columns of the selected inverse equal solves with unit
vectors on the factor's pattern."""
import numpy as np
import pytest

from densehelp import README_A, README_ORDERS, README_SELINV, dense_lower, permuted_dense, rel, to_upper_layout

pytestmark = pytest.mark.gpu


def _pattern(S):
    return dense_lower(S, np.ones(S.ndz), np.ones(S.nlz)) != 0


def _cases(ct):
    pr = ct.problems
    yield "lap2d_30_nd", pr.laplacian2d_dirichlet(30, 30), ct.grid_nd(30, 30, 1), "maximal"
    yield "lap3d_12_nd", pr.laplacian3d_7pt(12, 12, 12, seed=2), ct.grid_nd(12, 12, 12), "maximal"
    yield "st27_9_nd", pr.stencil3d_27pt(9, 9, 9, seed=3), ct.grid_nd(9, 9, 9), "maximal"
    yield "lap3d_10_mmd", pr.laplacian3d_7pt(10, 10, 10, seed=4), "mmd", "maximal"
    yield "lap3d_8_fund", pr.laplacian3d_7pt(8, 8, 8, seed=5), ct.grid_nd(8, 8, 8), "fundamental"
    yield "lap3d_20_nd", pr.laplacian3d_7pt(20, 20, 20, seed=6), ct.grid_nd(20, 20, 20), "maximal"
    yield "st27_14_nd", pr.stencil3d_27pt(14, 14, 14, seed=7), ct.grid_nd(14, 14, 14), "maximal"


CASES = ["lap2d_30_nd", "lap3d_12_nd", "st27_9_nd", "lap3d_10_mmd", "lap3d_8_fund", "lap3d_20_nd",
         "st27_14_nd"]


@pytest.fixture(scope="module")
def cases(ct):
    return {name: rest for name, *rest in _cases(ct)}


@pytest.mark.parametrize("order", README_ORDERS)
def test_selinv_docstring_on_device(ct, order):
    """selinv.jl:14-31: selinv!(cholesky!(ChordalCholesky(A))) prints this
    lower triangle; the entries are dyadic."""
    print(f"selinv docstring order={order}", flush=True)
    F = ct.ChordalCholesky(README_A, alg=order)
    ct.cholesky_(F)
    ct.selinv_(F)
    got = dense_lower(F.S, F.Dval, F.Lval)
    assert np.allclose(got, README_SELINV, rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", CASES)
def test_selinv_matches_oracle(ct, oracle, cases, name):
    A, order, snd = cases[name]
    print(f"selinv case {name}", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D, L, d, info = oracle.chol_impl(S, Dv, Lv)
    OD, OL = oracle.selinv(S, D, L)
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    ct.selinv_(F)
    SD, SL = F.Dval, F.Lval
    assert rel(SD, OD) <= 1e-12 and rel(SL, OL) <= 1e-12
    if A.shape[0] <= 3000:
        inv = np.linalg.inv(permuted_dense(A, P))
        mask = _pattern(S)
        assert rel(dense_lower(S, SD, SL)[mask], np.tril(inv)[mask]) <= 1e-12
    # the factor is kept: a solve after the inversion is the factor's
    B = np.random.default_rng(1).standard_normal(A.shape[0])
    assert rel(ct.solve(F, B), np.asarray(oracle.solve(S, P, D, L, d, B))) <= 1e-11
    # a second inversion of the same factor gives the same values bit for bit
    ct.selinv_(F)
    assert np.array_equal(F.Dval, SD) and np.array_equal(F.Lval, SL)


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_selinv_impl_dropin(ct, oracle, cases, uplo):
    """selinv_impl! on host arrays in either layout, twice on one plan."""
    A, order, snd = cases["lap3d_12_nd"]
    print(f"selinv_impl uplo={uplo}", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D, L, d, info = oracle.chol_impl(S, Dv, Lv)
    OD, OL = oracle.selinv(S, D, L)
    if uplo == "U":
        D, L = to_upper_layout(S, D, L)
        OD, OL = to_upper_layout(S, OD, OL)
    for _ in range(2):
        SD, SL = ct.selinv_impl(S, D, L, uplo=uplo)
        assert rel(SD, OD) <= 1e-12 and rel(SL, OL) <= 1e-12


def test_selinv_fp32_fronts(ct, oracle, cases):
    A, order, snd = cases["lap3d_20_nd"]
    print("selinv fp32", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D, L, d, info = oracle.chol_impl(S, Dv, Lv)
    OD, OL = oracle.selinv(S, D, L)
    F = ct.ChordalCholesky(A, P, S, precision=ct.MF_FP32)
    ct.cholesky_(F)
    ct.selinv_(F)
    # FP32 fronts: the error grows with the condition number (about 1e3 here)
    assert rel(F.Dval, OD) <= 2e-4 and rel(F.Lval, OL) <= 2e-4


def test_selinv_needs_cholesky(ct, cases):
    A, order, snd = cases["lap3d_8_fund"]
    F = ct.ChordalLDLt(A, alg=order, snd=snd)
    ct.ldlt_(F)
    with pytest.raises(TypeError):
        ct.selinv_(F)
    from cliquetrees_amd import _native
    rc = _native.hip().mf_selinv(F._f.handle)
    assert rc == -1 and b"Cholesky" in _native.hip().mf_last_error()


def test_selinv_lap7_64_columns_are_solves(ct):
    """64^3 7-point Laplacian: for a few permuted columns j the selected
    inverse's column j (its stored rows) equals the solve A x = e_perm[j]
    read at those rows; the diagonal is positive."""
    A = ct.problems.laplacian3d_7pt(64, 64, 64, seed=2)
    print("selinv lap7_64", flush=True)
    F = ct.ChordalCholesky(A, alg=ct.grid_nd(64, 64, 64))
    ct.cholesky_(F)
    ct.selinv_(F)
    S, P = F.S, F.P
    SD, SL = F.Dval, F.Lval
    rng = np.random.default_rng(11)
    fronts = rng.choice(S.nfr, size=4, replace=False).tolist() + [S.nfr - 1]
    for j in fronts:
        r0 = int(S.resptr[j] - 1)
        nn, na = int(S.resptr[j + 1] - S.resptr[j]), int(S.sepptr[j + 1] - S.sepptr[j])
        c = 0
        e = np.zeros(S.n)
        e[P.perm[r0 + c] - 1] = 1.0
        x = ct.solve(F, e)[P.perm - 1]  # permuted order
        col_d = SD[S.Dptr[j] - 1 + c * nn: S.Dptr[j] - 1 + (c + 1) * nn][c:]
        rows = S.septgt[S.sepptr[j] - 1: S.sepptr[j] - 1 + na] - 1
        col_l = SL[S.Lptr[j] - 1 + c * na: S.Lptr[j] - 1 + (c + 1) * na]
        want = np.concatenate([x[r0 + c: r0 + nn], x[rows]])
        got = np.concatenate([col_d, col_l])
        assert rel(got, want) <= 1e-10, j
        assert col_d[0] > 0
