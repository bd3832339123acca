"""Pins of the CPU oracle (oracle/oracle.c) that do not go through the oracle's
own formulas: the factors and solutions printed in the reference's
documentation, dense LAPACK on PAP' (the Cholesky factor of an SPD matrix and
the unit LDL' factor of a quasi-definite matrix are unique), leading
principal minors for the failing column, and the defining properties of the
regularization rules (no_regularization.jl:27-42,
dynamic_regularization.jl:75-90).  CPU only: the GPU tests then compare the
HIP engine with this pinned oracle and with the same known answers.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from densehelp import (DYNREG_A, DYNREG_B, README_A, README_AI, README_D, README_L, README_LU,
                       README_ORDERS, README_PERM, README_SELINV, dense_lower, first_indefinite_minor,
                       leading_pivots, nod_columns, outside_pattern_norm, permuted_dense, rel,
                       to_layout, weakly_indefinite)


def _chol(ct, oracle, A, alg=None, diag=0, signs=None, reg=None, snd="maximal"):
    P, S = ct.symbolic(A, alg=alg, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D, L, d, info = oracle.chol_impl(S, Dv, Lv, diag=diag, signs=signs, reg=reg)
    return P, S, D, L, d, info


def _dyn(ct):
    return ct.DynamicRegularization().initialize(np.finfo(float).eps)


# ---- README known answers --------------------------------------------------------

@pytest.mark.parametrize("order", README_ORDERS)
def test_readme_cholesky_printed_factor(ct, oracle, order):
    """src/Multifrontal.jl/README.md:29-44: perm [3,1,2,4,5] and L as printed;
    F \\ [4,2,0,0,2] = e1 (:50-58)."""
    P, S, D, L, d, info = _chol(ct, oracle, README_A, alg=order)
    assert info == 0 and P.perm.tolist() == README_PERM
    assert np.array_equal(dense_lower(S, D, L), README_L)
    x = oracle.solve(S, P, D, L, d, np.array([4.0, 2, 0, 0, 2]))
    assert np.allclose(x, [1, 0, 0, 0, 0], rtol=0, atol=1e-15)


@pytest.mark.parametrize("order", README_ORDERS)
def test_readme_ldlt_printed_factor(ct, oracle, order):
    """README.md:89-124: unit L and D = diag(4,-4,-4,4,8) as printed;
    F \\ [-4,-2,0,0,2] = e1."""
    P, S, D, L, d, info = _chol(ct, oracle, README_AI, alg=order, diag=1)
    assert info == 0 and P.perm.tolist() == README_PERM
    assert np.array_equal(dense_lower(S, D, L, unit=True), README_LU)
    assert np.array_equal(d, README_D)
    x = oracle.solve(S, P, D, L, d, np.array([-4.0, -2, 0, 0, 2]), diag=1)
    assert np.allclose(x, [1, 0, 0, 0, 0], rtol=0, atol=1e-15)


def test_readme_ldlt_signs(ct, oracle):
    """README.md:145-157: signs = [-1,-1,1,1,1] (given in the original order and
    permuted into factor order, utils.jl:1-13) leave the printed factor; signs
    that disagree with a pivot fail at the first such column in factor order
    (no_regularization.jl:27-33)."""
    order = README_ORDERS[0]
    P, S = ct.symbolic(README_AI, alg=order)
    Dv, Lv = oracle.assemble(S, P, README_AI)
    s = np.array([-1.0, -1, 1, 1, 1])[P.perm - 1]
    D, L, d, info = oracle.chol_impl(S, Dv, Lv, diag=1, signs=s)
    assert info == 0 and np.array_equal(d, README_D)
    assert np.array_equal(dense_lower(S, D, L, unit=True), README_LU)
    # closed form: the pivots in factor order are README_D; the first k whose
    # sign disagrees is the failing column
    for orig in ([1.0, 1, 1, 1, 1], [-1.0, -1, 1, 1, -1], [-1.0, 1, 1, 1, 1], [-1.0, -1, -1, 1, 1]):
        sp_ = np.array(orig)[P.perm - 1]
        bad = np.nonzero(np.sign(README_D) != sp_)[0]
        expect = int(bad[0]) + 1 if len(bad) else 0
        _, _, _, info = oracle.chol_impl(S, Dv, Lv, diag=1, signs=sp_)
        assert info == expect, (orig, info, expect)
    # zero signs mean "no check"
    _, _, d0, info = oracle.chol_impl(S, Dv, Lv, diag=1, signs=np.zeros(5))
    assert info == 0 and np.array_equal(d0, README_D)


def test_dynamic_regularization_docstring(ct, oracle):
    """dynamic_regularization.jl:12-30: Matrix(F) = P' L L' P as printed (six
    significant digits) with delta = cbrt(eps), epsilon = sqrt(eps)."""
    P, S, D, L, d, info = _chol(ct, oracle, DYNREG_A, signs=np.ones(3), reg=_dyn(ct))
    assert info == 0
    Ld = dense_lower(S, D, L)
    B = np.empty((3, 3))
    p = P.perm - 1
    B[np.ix_(p, p)] = Ld @ Ld.T
    printed = np.vectorize(lambda x: float(f"{x:.6g}"))(B)
    assert np.array_equal(printed, DYNREG_B)
    np.linalg.cholesky(B)          # isposdef(B) is true: a Cholesky factor exists


# ---- dense LAPACK on PAP' --------------------------------------------------------

def _spd_cases(ct):
    pr = ct.problems
    yield "lap2d_12_nd", pr.laplacian2d_dirichlet(12, 12), ct.grid_nd(12, 12, 1), "maximal"
    yield "lap2d_9_nat", pr.laplacian2d_dirichlet(9, 9), None, "maximal"
    yield "lap3d_8_nd", pr.laplacian3d_7pt(8, 8, 8, seed=2), ct.grid_nd(8, 8, 8), "maximal"
    yield "lap3d_7_nodal", pr.laplacian3d_7pt(7, 6, 5, seed=3), ct.grid_nd(7, 6, 5), "nodal"
    yield "st27_6_fund", pr.stencil3d_27pt(6, 6, 6, seed=3), ct.grid_nd(6, 6, 6), "fundamental"
    yield "lap3d_10_mmd", pr.laplacian3d_7pt(10, 10, 10, seed=4), "mmd", "maximal"
    rng = np.random.default_rng(7)
    R = sp.random(300, 300, density=0.02, random_state=rng)
    A = (R + R.T + 10.0 * sp.identity(300)).tocsc()
    yield "random300_mmd", A, "mmd", "maximal"


SPD = ["lap2d_12_nd", "lap2d_9_nat", "lap3d_8_nd", "lap3d_7_nodal", "st27_6_fund", "lap3d_10_mmd",
       "random300_mmd"]


@pytest.fixture(scope="module")
def spd(ct):
    return {name: rest for name, *rest in _spd_cases(ct)}


@pytest.mark.parametrize("name", SPD)
def test_cholesky_vs_dense_lapack(ct, oracle, spd, name):
    """numpy.linalg.cholesky(PAP') mapped into the Dval/Lval layout equals the
    oracle's factor to 1e-13; the dense factor has nothing outside the
    symbolic pattern (the pattern is the fill of PAP')."""
    A, order, snd = spd[name]
    P, S, D, L, d, info = _chol(ct, oracle, A, alg=order, snd=snd)
    assert info == 0
    Ld = np.linalg.cholesky(permuted_dense(A, P))
    assert outside_pattern_norm(S, Ld) <= 1e-13 * np.linalg.norm(Ld)
    D1, L1 = to_layout(S, Ld)
    assert rel(D, D1) <= 1e-13 and rel(L, L1) <= 1e-13


@pytest.mark.parametrize("name", SPD)
def test_ldlt_vs_dense_lapack(ct, oracle, spd, name):
    """On an SPD matrix the unit LDL' factor is L_chol diag(L_chol)^-1 with
    D = diag(L_chol)^2; the Dval diagonal holds the pivot (nn > 1, potrf.jl:5-10)
    or the assembled entry of PAP' (nn = 1, cholesky.jl:643-659)."""
    A, order, snd = spd[name]
    P, S, D, L, d, info = _chol(ct, oracle, A, alg=order, diag=1, snd=snd)
    assert info == 0
    M = permuted_dense(A, P)
    Lc = np.linalg.cholesky(M)
    piv = np.diag(Lc) ** 2
    Lu = Lc / np.diag(Lc)[None, :]
    assert rel(d, piv) <= 1e-13
    D1, L1 = to_layout(S, Lu)
    diag_expect = piv.copy()
    nod = nod_columns(S)
    diag_expect[nod] = np.diag(M)[nod]
    np.fill_diagonal(Lu, 0.0)
    Lu += np.diag(diag_expect)
    D1, L1 = to_layout(S, Lu)
    assert rel(L, L1) <= 1e-13 and rel(D, D1) <= 1e-13


@pytest.mark.parametrize("name", SPD)
def test_solve_vs_dense(ct, oracle, spd, name):
    """ldiv! through the oracle equals a dense solve with A, and the natural
    solve equals triangular solves with the dense factor of PAP'."""
    A, order, snd = spd[name]
    P, S, D, L, d, info = _chol(ct, oracle, A, alg=order, snd=snd)
    B = np.random.default_rng(3).standard_normal((A.shape[0], 5))
    X = oracle.solve(S, P, D, L, d, B)
    X1 = np.linalg.solve(A.toarray(), B)
    assert rel(X, X1) <= 1e-12
    Ld = dense_lower(S, D, L)
    C = P.apply(B)
    Y = sla.solve_triangular(Ld.T, sla.solve_triangular(Ld, C, lower=True), lower=False)
    assert rel(oracle.div_natural(S, D, L, d, C), Y) <= 1e-13


def test_ldlt_quasidefinite_reconstruction(ct, oracle):
    """Quasi-definite KKT matrix [H A'; A -1e-6 I]: the unit LDL' without
    pivoting is unique; L D L' = PAP' to rounding, L is inside the symbolic
    pattern, and D equals the ratios of leading principal minors of PAP'."""
    K, nh, m = ct.problems.kkt_matrix((5, 4, 4), seed=5)
    P, S, D, L, d, info = _chol(ct, oracle, K, alg="mmd", diag=1)
    assert info == 0
    M = permuted_dense(K, P)
    Lu = dense_lower(S, D, L, unit=True)
    R = Lu @ np.diag(d) @ Lu.T
    # backward error of an unpivoted LDL' is bounded by eps |L| |D| |L'|
    # (the pivots of this ordering span 1e-6 .. 1e6)
    G = np.abs(Lu) @ np.diag(np.abs(d)) @ np.abs(Lu).T
    assert np.abs(R - M).max() <= 1e-13 * G.max()
    assert rel(d, leading_pivots(M)) <= 1e-9
    # the signs of the pivots follow the blocks (quasi-definite)
    signs = np.concatenate([np.ones(nh), -np.ones(m)])[P.perm - 1]
    assert np.array_equal(np.sign(d), signs)


# ---- failures and regularization ---------------------------------------------

def _indefinite_cases(ct):
    pr = ct.problems
    yield pr.laplacian2d_dirichlet(10, 10) - 3.5 * sp.identity(100), ct.grid_nd(10, 10, 1)
    yield pr.laplacian2d_dirichlet(10, 10) - 1.0 * sp.identity(100), None
    yield pr.laplacian3d_7pt(5, 5, 5, seed=2) - 0.4 * sp.identity(125), ct.grid_nd(5, 5, 5)
    yield pr.laplacian3d_7pt(6, 5, 4, seed=8) - 1.5 * sp.identity(120), "mmd"


def test_failure_info_is_first_indefinite_leading_minor(ct, oracle):
    """chol_impl! returns the first failing column in permuted order
    (cholesky.jl:384-388): the first k whose leading k x k block of PAP' is not
    positive definite."""
    for A, order in _indefinite_cases(ct):
        A = sp.csc_matrix(A)
        P, S, D, L, d, info = _chol(ct, oracle, A, alg=order)
        k = first_indefinite_minor(permuted_dense(A, P))
        assert k > 0 and info == k, (info, k)


def test_ldlt_sign_mismatch_info(ct, oracle):
    """ldlt! with signs = +1 everywhere on an indefinite matrix fails at the
    first pivot whose sign is negative; the pivots come from leading minors."""
    for A, order in _indefinite_cases(ct):
        A = sp.csc_matrix(A)
        P, S = ct.symbolic(A, alg=order)
        Dv, Lv = oracle.assemble(S, P, A)
        _, _, _, info = oracle.chol_impl(S, Dv, Lv, diag=1, signs=np.ones(S.n))
        piv = leading_pivots(permuted_dense(A, P))
        assert info == int(np.nonzero(piv <= 0)[0][0]) + 1


def _regularization_cases(ct):
    pr = ct.problems
    yield weakly_indefinite(pr.laplacian2d_dirichlet(10, 10), [3, 41, 77], 1), ct.grid_nd(10, 10, 1)
    yield weakly_indefinite(pr.laplacian3d_7pt(5, 5, 5, seed=2), [0, 62, 124, 99], 2), ct.grid_nd(5, 5, 5)
    yield weakly_indefinite(pr.laplacian3d_7pt(6, 5, 4, seed=8), [7, 8, 50], 3), "mmd"


@pytest.mark.parametrize("diag", [0, 1])
def test_dynamic_regularization_rule(ct, oracle, diag):
    """DynamicRegularization (dynamic_regularization.jl:75-90) on an indefinite
    matrix with signs = +1: the factor is exact for PAP' + E with E diagonal;
    at a replaced pivot the pivot is delta*s and the pivot before replacement
    satisfied s*Djj < epsilon, elsewhere E = 0 and s*Djj >= epsilon.  For LDL'
    the Dval diagonal keeps the pivot before replacement (potrf.jl:5-10) or,
    on nn = 1 fronts, the assembled entry (cholesky.jl:643-659)."""
    delta, eps = _dyn(ct)
    for A, order in _regularization_cases(ct):
        P, S, D, L, d, info = _chol(ct, oracle, A, alg=order, diag=diag, signs=np.ones(A.shape[0]),
                                    reg=(delta, eps))
        assert info == 0
        M = permuted_dense(A, P)
        if diag == 0:
            Ld = dense_lower(S, D, L)
            piv = np.diag(Ld) ** 2
            R = Ld @ Ld.T
        else:
            Lu = dense_lower(S, D, L, unit=True)
            piv = d
            R = Lu @ np.diag(d) @ Lu.T
        E = R - M
        off = E - np.diag(np.diag(E))
        scale = np.abs(M).max() + np.abs(np.diag(E)).max()
        assert np.abs(off).max() <= 1e-11 * scale
        e = np.diag(E)
        hit = np.abs(e) > 1e-9 * scale
        assert hit.any() and not hit.all()
        assert np.allclose(piv[hit], delta, rtol=1e-12, atol=0)
        before = piv - e
        assert np.all(before[hit] < eps) and np.all(piv[~hit] >= eps)
        if diag == 1:
            Ddiag = dense_lower(S, D, L).diagonal()
            nod = np.zeros(S.n, dtype=bool)
            nod[nod_columns(S)] = True
            assert np.allclose(Ddiag[nod], np.diag(M)[nod], rtol=0, atol=0)
            assert np.allclose(Ddiag[~nod], before[~nod], rtol=1e-10, atol=1e-10 * scale)


def test_config1_lap2d_100_mmd_residual(ct, oracle):
    """The CPU config: 2-D 5-point Dirichlet Laplacian 100 x 100, MMD order,
    one N(0,1) right-hand side (seed 1): relative residual at rounding."""
    A = ct.problems.laplacian2d_dirichlet(100, 100)
    P, S, D, L, d, info = _chol(ct, oracle, A, alg="mmd")
    assert info == 0
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x = oracle.solve(S, P, D, L, d, b)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-14


@pytest.mark.parametrize("diag", [0, 1])
def test_host_blas_mode_matches_loops(ct, oracle, diag):
    """The oracle's host-BLAS mode (used for wide fronts and the CPU baseline)
    gives the factor of its C loops to rounding, with the strict upper
    triangles of the diagonal blocks left at zero (a front of 400 columns:
    several 128-column blocks of the blocked updates)."""
    A = ct.problems.laplacian3d_7pt(20, 20, 16, seed=6)
    P, S = ct.symbolic(A, alg=ct.grid_nd(20, 20, 16))
    assert S.nn().max() >= 300
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, _ = oracle.chol_impl(S, Dv, Lv, diag=diag)
    oracle.use_host_blas(True)
    try:
        D1, L1, d1, info = oracle.chol_impl(S, Dv, Lv, diag=diag)
    finally:
        oracle.use_host_blas(False)
    assert info == 0
    assert rel(D1, D0) <= 1e-13 and rel(L1, L0) <= 1e-13 and rel(d1, d0) <= 1e-13
    for j in range(S.nfr):
        k = S.nn()[j]
        blk = D1[S.Dptr[j] - 1: S.Dptr[j] - 1 + k * k].reshape(k, k, order="F")
        assert not np.triu(blk, 1).any()


# ---- selected inversion (selinv.jl:38-424) --------------------------------------

def _pattern(S):
    return dense_lower(S, np.ones(S.ndz), np.ones(S.nlz)) != 0


@pytest.mark.parametrize("order", README_ORDERS)
def test_selinv_docstring(ct, oracle, order):
    """selinv.jl:14-31: the printed selected inverse of the README matrix
    (10 stored entries, perm [3,1,2,4,5]); its entries are dyadic, so exact."""
    P, S, D, L, d, info = _chol(ct, oracle, README_A, alg=order)
    Ds, Ls = oracle.selinv(S, D, L)
    assert int(_pattern(S).sum()) == 10
    assert np.array_equal(dense_lower(S, Ds, Ls), README_SELINV)


@pytest.mark.parametrize("name", SPD)
def test_selinv_vs_dense_inverse(ct, oracle, spd, name):
    """The selected inverse equals numpy.linalg.inv(PAP') on every entry of the
    factor's pattern (lower triangle) to 1e-12, the strict upper triangle of
    each diagonal block is 0 (copyrec! of F11, selinv.jl:250-255), and nothing
    outside the pattern is written."""
    A, order, snd = spd[name]
    P, S, D, L, d, info = _chol(ct, oracle, A, alg=order, snd=snd)
    Ds, Ls = oracle.selinv(S, D, L)
    inv = np.linalg.inv(permuted_dense(A, P))
    mask = _pattern(S)
    got = dense_lower(S, Ds, Ls)
    assert rel(got[mask], np.tril(inv)[mask]) <= 1e-12
    for j in range(S.nfr):
        k = int(S.nn()[j])
        blk = Ds[S.Dptr[j] - 1: S.Dptr[j] - 1 + k * k].reshape(k, k, order="F")
        assert not np.triu(blk, 1).any()


def test_selinv_diagonal_and_chain(ct, oracle):
    """Closed forms: a diagonal matrix inverts entrywise; the path graph's
    tridiagonal Laplacian + I has a known inverse on its (tridiagonal) pattern."""
    a = np.array([2.0, 4.0, 8.0, 0.5])
    P, S, D, L, d, info = _chol(ct, oracle, sp.diags(a).tocsc())
    Ds, Ls = oracle.selinv(S, D, L)
    got = np.diag(dense_lower(S, Ds, Ls))
    assert np.allclose(got, 1.0 / a[P.perm - 1], rtol=1e-15)
    n = 40
    T = sp.diags([-np.ones(n - 1), 3.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsc()
    P, S, D, L, d, info = _chol(ct, oracle, T, alg=list(range(1, n + 1)))
    Ds, Ls = oracle.selinv(S, D, L)
    inv = np.linalg.inv(T.toarray())
    got = dense_lower(S, Ds, Ls)
    mask = _pattern(S)
    assert rel(got[mask], np.tril(inv)[mask]) <= 1e-14


# ---- completion (complete.jl:1-451) ------------------------------------------

@pytest.mark.parametrize("name", SPD)
def test_complete_inverts_selinv(ct, oracle, spd, name):
    """The inverse of PAP' is the maximum-determinant completion of its own
    entries on the chordal pattern (its inverse PAP' vanishes off the
    pattern), so complete!(selinv!(cholesky!(A))) is cholesky!(A) again."""
    A, order, snd = spd[name]
    P, S, D, L, d, info = _chol(ct, oracle, A, alg=order, snd=snd)
    SD, SL = oracle.selinv(S, D, L)
    CD, CL, info = oracle.complete(S, SD, SL)
    assert info == 0
    assert rel(dense_lower(S, CD, CL), dense_lower(S, D, L)) <= 1e-11


def test_complete_of_a_dense_pattern(ct, oracle):
    """On a dense 6 x 6 pattern (one front) the completion is the matrix
    itself: the result is the Cholesky factor of its inverse."""
    rng = np.random.default_rng(3)
    G = rng.standard_normal((6, 6))
    Sig = G @ G.T + 6 * np.eye(6)
    P, S = ct.symbolic(sp.csc_matrix(np.ones((6, 6))))
    Sp = permuted_dense(Sig, P)
    Dv, Lv = to_layout(S, np.tril(Sp))
    CD, CL, info = oracle.complete(S, Dv, Lv)
    assert info == 0
    want = np.linalg.cholesky(np.linalg.inv(Sp))
    assert rel(dense_lower(S, CD, CL), want) <= 1e-12


def test_complete_failure_info(ct, oracle):
    """An indefinite separator block stops the completion with its potrf!
    failure (a positive info)."""
    A = ct.problems.laplacian2d_dirichlet(8, 8)
    P, S, D, L, d, info = _chol(ct, oracle, A, alg=ct.grid_nd(8, 8, 1))
    SD, SL = oracle.selinv(S, D, L)
    CD, CL, info = oracle.complete(S, -SD, -SL)
    assert info > 0
