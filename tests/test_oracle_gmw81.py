"""The oracle's GMW81 rule (regularization/gmw81.jl, chol_factor!'s GMW81
form cholesky.jl:739-875, the nn = 1 form of chol_loop_nod!) pinned by:

* the docstrings' printed results: GMW81's own example (gmw81.jl:8-30,
  Matrix(F)) and the reference test's pivots diag(F) = [3.771, 5.750, 1.121]
  (test/multifrontal.jl:190-202, 242-260);
* a dense restatement of the published algorithm (Gill, Murray and Wright,
  Practical Optimization, 1981, sec. 4.4.2.2, in the reference's form
  without the diagonal interchange): on PAP' each pivot becomes
  s_j max(|c_jj|, theta_j^2 / beta, delta) with theta_j the largest
  magnitude below it in the current Schur complement; the chordal pattern
  holds every nonzero of the Schur complement, so the multifrontal result
  is the dense one;
* properties the method fixes: L L' - PAP' (L D L' - PAP') is diagonal with
  s_j E_jj >= 0, and a positive definite matrix whose pivots stay above
  delta is not modified (theta_j^2 <= c_ii c_jj <= beta c_jj).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from densehelp import dense_lower, permuted_dense, rel

EPS = np.finfo(np.float64).eps
GMW_A = np.array([[1, 1, 2], [1, 1, 3], [2, 3, 1]], float)
GMW_B = np.array([[3.77124, 1.0, 2.0], [1.0, 6.01561, 3.0], [2.0, 3.0, 3.24264]])
GMW_D = np.array([3.771, 5.750, 1.121])


def beta_of(A):
    """gmw81_beta (gmw81.jl:90-94) from the matrix itself."""
    Ad = A.toarray() if sp.issparse(A) else np.asarray(A, float)
    n = Ad.shape[0]
    gamma = np.abs(np.diag(Ad)).max()
    xi = np.abs(np.tril(Ad, -1)).max() if n > 1 else 0.0
    return max(gamma, xi / np.sqrt(n * n - 1.0), EPS)


def gmw81_dense(M, s, beta, delta, ldlt):
    """Right-looking GMW81 on the dense symmetric M: returns (L, d), L unit
    lower with d for LDLt, L with the pivots' square roots for Cholesky."""
    C = np.array(M, dtype=np.float64, copy=True)
    n = C.shape[0]
    L = np.zeros((n, n))
    d = np.zeros(n)
    for j in range(n):
        theta = np.abs(C[j + 1:, j]).max() if j + 1 < n else 0.0
        p = s[j] * max(abs(C[j, j]), theta * theta / beta, delta)
        d[j] = p
        if ldlt:
            L[j, j] = 1.0
            L[j + 1:, j] = C[j + 1:, j] / p
            C[j + 1:, j + 1:] -= np.outer(L[j + 1:, j], L[j + 1:, j]) * p
        else:
            r = np.sqrt(p)
            L[j, j] = r
            L[j + 1:, j] = C[j + 1:, j] / r
            C[j + 1:, j + 1:] -= np.outer(L[j + 1:, j], L[j + 1:, j])
    return L, d


def _shifted(A, frac):
    """A - frac * lambda_max(A) I: indefinite for frac inside the spectrum."""
    lmax = sp.linalg.eigsh(sp.csc_matrix(A, dtype=np.float64), k=1, which="LA",
                           return_eigenvectors=False)[0]
    return sp.csc_matrix(A - frac * lmax * sp.eye(A.shape[0]))


def _cases(ct):
    pr = ct.problems
    yield "lap2d_20_mmd", _shifted(pr.laplacian2d_dirichlet(20, 20), 0.3), "mmd", "maximal"
    yield "lap2d_16_fund", _shifted(pr.laplacian2d_dirichlet(16, 16), 0.1), "mmd", "fundamental"
    yield "lap3d_10_nd", _shifted(pr.laplacian3d_7pt(10, 10, 10, seed=2), 0.2), ct.grid_nd(10, 10, 10), "maximal"
    yield "st27_8_nd", _shifted(pr.stencil3d_27pt(8, 8, 8, seed=3), 0.05), ct.grid_nd(8, 8, 8), "maximal"


CASES = ["lap2d_20_mmd", "lap2d_16_fund", "lap3d_10_nd", "st27_8_nd"]


@pytest.fixture(scope="module")
def cases(ct):
    return {name: rest for name, *rest in _cases(ct)}


def _gmw(oracle, ct, A, order, snd="maximal", diag=0, signs=None, beta=-1.0, delta=-1.0):
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    if signs is not None:
        signs = np.ascontiguousarray(np.asarray(signs, float)[P.perm - 1])
    D, L, d, info = oracle.chol_impl(S, Dv, Lv, diag, signs=signs, reg=("gmw81", beta, delta))
    return P, S, D, L, d, info


@pytest.mark.parametrize("order", [[1, 2, 3], [3, 1, 2], [2, 3, 1]])
def test_gmw81_docstring(ct, oracle, order):
    """gmw81.jl:8-30: Matrix(F) as printed (six significant digits), and the
    reference test's pivots for Cholesky and LDLt (signs = ones)."""
    A = sp.csc_matrix(GMW_A)
    # GMW81's result depends on the elimination order: every order is
    # compared with the dense restatement in its permuted frame, the
    # identity order also with the printed values
    P, S, D, L, d, info = _gmw(oracle, ct, A, np.array(order))
    assert info == 0
    Lf = dense_lower(S, D, L)
    ip = np.argsort(P.perm - 1)
    B = (Lf @ Lf.T)[np.ix_(ip, ip)]
    M = permuted_dense(A, P)
    Ld, dd = gmw81_dense(M, np.ones(3), beta_of(A), np.cbrt(EPS), False)
    assert np.allclose(Lf, Ld, rtol=0, atol=1e-14)
    if order == [1, 2, 3]:
        assert np.allclose(B, GMW_B, rtol=0, atol=6e-6)
        assert np.allclose(dd, GMW_D, rtol=1e-3)
    P2, S2, D2, L2, d2, info2 = _gmw(oracle, ct, A, np.array(order), diag=1, signs=np.ones(3))
    assert info2 == 0
    assert np.allclose(d2, dd, rtol=0, atol=1e-14)
    if order == [1, 2, 3]:
        assert np.allclose(d2, GMW_D, rtol=1e-3)


def test_gmw81_beta(ct, oracle, cases):
    for name in CASES:
        A, order, snd = cases[name]
        P, S = ct.symbolic(A, alg=order, snd=snd)
        Dv, Lv = oracle.assemble(S, P, A)
        assert oracle.gmw81_beta(S, Dv, Lv) == pytest.approx(beta_of(A), rel=1e-15)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("diag", [0, 1])
def test_gmw81_matches_dense(ct, oracle, cases, name, diag):
    A, order, snd = cases[name]
    n = A.shape[0]
    signs = np.ones(n) if diag else None
    P, S, D, L, d, info = _gmw(oracle, ct, A, order, snd, diag, signs)
    assert info == 0
    M = permuted_dense(A, P)
    Ld, dd = gmw81_dense(M, np.ones(n), beta_of(A), np.cbrt(EPS), diag == 1)
    Lf = dense_lower(S, D, L, unit=diag == 1)
    assert rel(Lf, Ld) <= 1e-12
    if diag:
        assert rel(d, dd) <= 1e-12
    # some pivots were modified, or the case would pin nothing
    E = ((Lf * d) @ Lf.T if diag else Lf @ Lf.T) - M
    assert np.abs(np.diag(E)).max() > 1e-3


@pytest.mark.parametrize("name", CASES)
def test_gmw81_perturbation_is_diagonal(ct, oracle, cases, name):
    """E = L L' - PAP' is diagonal and E >= 0 (GMW81 only raises pivots)."""
    A, order, snd = cases[name]
    P, S, D, L, d, info = _gmw(oracle, ct, A, order, snd)
    Lf = dense_lower(S, D, L)
    M = permuted_dense(A, P)
    E = Lf @ Lf.T - M
    off = E - np.diag(np.diag(E))
    assert np.abs(off).max() <= 1e-11 * np.abs(M).max()
    assert np.diag(E).min() >= -1e-11 * np.abs(M).max()


def test_gmw81_ldlt_kkt_signs(ct, oracle):
    """Quasi-definite KKT [H A'; A -1e-6 I] with signs (+1 on H, -1 on the
    multipliers): LDLt with GMW81 equals the dense restatement, and the
    perturbation is diagonal with s_j E_jj >= 0."""
    K, nh, m = ct.problems.kkt_matrix(ct.problems.near_cubic_dims(300, 3), seed=7)
    n = K.shape[0]
    signs = np.concatenate([np.ones(nh), -np.ones(n - nh)])
    P, S, D, L, d, info = _gmw(oracle, ct, K, "mmd", diag=1, signs=signs)
    assert info == 0
    M = permuted_dense(K, P)
    sp_ = signs[P.perm - 1]
    Ld, dd = gmw81_dense(M, sp_, beta_of(K), np.cbrt(EPS), True)
    Lf = dense_lower(S, D, L, unit=True)
    assert rel(Lf, Ld) <= 1e-12 and rel(d, dd) <= 1e-12
    E = (Lf * d) @ Lf.T - M
    off = E - np.diag(np.diag(E))
    assert np.abs(off).max() <= 1e-10 * np.abs(M).max()
    assert (sp_ * np.diag(E)).min() >= -1e-10 * np.abs(M).max()


def test_gmw81_spd_unmodified(ct, oracle):
    """A positive definite matrix with pivots above delta is left as it is:
    the GMW81 factor equals the plain Cholesky factor."""
    A = ct.problems.laplacian3d_7pt(9, 9, 9, seed=4)
    order = ct.grid_nd(9, 9, 9)
    P, S, D, L, d, info = _gmw(oracle, ct, A, order)
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv, 0)
    assert info == 0 and info0 == 0
    assert rel(D, D0) <= 1e-14 and rel(L, L0) <= 1e-14


def test_gmw81_parameters(ct, oracle, cases):
    """Given beta and delta are used as they are: a delta above every pivot
    makes every pivot delta (Cholesky: L's diagonal sqrt(delta) on the first
    column)."""
    A, order, snd = cases["lap2d_20_mmd"]
    n = A.shape[0]
    P, S, D, L, d, info = _gmw(oracle, ct, A, order, snd, beta=2.5, delta=0.75)
    M = permuted_dense(A, P)
    Ld, dd = gmw81_dense(M, np.ones(n), 2.5, 0.75, False)
    assert rel(dense_lower(S, D, L), Ld) <= 1e-12
