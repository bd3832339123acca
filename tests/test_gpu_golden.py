"""The HIP engine against the known answers that pin the oracle: the factors
printed in the reference's README and DynamicRegularization docstring, dense
LAPACK on PAP', leading principal minors for the failing column and the
regularization rule, each checked on the device result directly and against
the oracle.  Also fronts wide enough (more than ten 64-column panel blocks)
to reach the grouped trailing updates of the tile schedule.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from densehelp import (DYNREG_A, DYNREG_B, README_A, README_AI, README_D, README_L, README_LU,
                       README_ORDERS, README_PERM, dense_lower, first_indefinite_minor,
                       leading_pivots, nod_columns, permuted_dense, rel, to_layout,
                       weakly_indefinite)

pytestmark = pytest.mark.gpu


def _dyn(ct):
    return ct.DynamicRegularization().initialize(np.finfo(float).eps)


@pytest.mark.parametrize("order", README_ORDERS)
def test_readme_cholesky_golden(ct, order):
    """src/Multifrontal.jl/README.md:29-58 on the device: perm, L as printed,
    F \\ b = e1, through the API and through the chol_impl! seam."""
    print(f"readme cholesky order={order}", flush=True)
    F = ct.ChordalCholesky(README_A, alg=order)
    ct.cholesky_(F)
    assert F.perm.tolist() == README_PERM and ct.issuccess(F)
    assert np.allclose(dense_lower(F.S, F.Dval, F.Lval), README_L, rtol=0, atol=1e-15)
    x = ct.solve(F, np.array([4.0, 2, 0, 0, 2]))
    assert np.allclose(x, [1, 0, 0, 0, 0], rtol=0, atol=1e-15)
    P, S = F.P, F.S
    Dv, Lv = to_layout(S, np.tril(permuted_dense(README_A, P)))
    D1, L1, _, info = ct.chol_impl(S, Dv, Lv, diag=ct.MF_CHOLESKY)
    assert info == 0 and np.allclose(dense_lower(S, D1, L1), README_L, rtol=0, atol=1e-15)


@pytest.mark.parametrize("order", README_ORDERS)
def test_readme_ldlt_golden(ct, order):
    """README.md:89-157 on the device: unit L and D as printed, the signs
    variant, and F \\ b = e1."""
    print(f"readme ldlt order={order}", flush=True)
    for signs in (None, [-1, -1, 1, 1, 1]):
        F = ct.ChordalLDLt(README_AI, alg=order)
        ct.ldlt_(F, signs=signs)
        assert F.perm.tolist() == README_PERM and ct.issuccess(F)
        assert np.allclose(dense_lower(F.S, F.Dval, F.Lval, unit=True), README_LU, rtol=0, atol=1e-15)
        assert np.allclose(F.d, README_D, rtol=0, atol=1e-14)
        x = ct.solve(F, np.array([-4.0, -2, 0, 0, 2]))
        assert np.allclose(x, [1, 0, 0, 0, 0], rtol=0, atol=1e-15)


def test_readme_ldlt_sign_mismatch(ct, oracle):
    """Signs that disagree with a pivot (no_regularization.jl:27-33): info is the
    first disagreeing column in factor order (closed form from the printed D),
    F.info maps it through perm (cholesky.jl:250-254), check raises
    SingularException with the permuted column (utils.jl:775-785)."""
    order = README_ORDERS[0]
    for orig in ([1.0, 1, 1, 1, 1], [-1.0, -1, 1, 1, -1], [-1.0, 1, 1, 1, 1]):
        F = ct.ChordalLDLt(README_AI, alg=order)
        s = np.array(orig)[F.perm - 1]
        k = int(np.nonzero(np.sign(README_D) != s)[0][0]) + 1
        ct.ldlt_(F, signs=orig, check=False)
        assert not ct.issuccess(F) and F.info == F.perm[k - 1]
        Dv, Lv = oracle.assemble(F.S, F.P, sp.csc_matrix(README_AI))
        assert oracle.chol_impl(F.S, Dv, Lv, diag=1, signs=s)[3] == k
        G = ct.ChordalLDLt(README_AI, alg=order)
        with pytest.raises(ct.SingularException) as e:
            ct.ldlt_(G, signs=orig)
        assert e.value.info == k


def test_dynamic_regularization_golden(ct):
    """dynamic_regularization.jl:12-30 on the device: Matrix(F) as printed."""
    F = ct.ChordalCholesky(DYNREG_A)
    ct.cholesky_(F, reg=ct.DynamicRegularization())
    assert ct.issuccess(F)
    Ld = dense_lower(F.S, F.Dval, F.Lval)
    B = np.empty((3, 3))
    p = F.perm - 1
    B[np.ix_(p, p)] = Ld @ Ld.T
    printed = np.vectorize(lambda x: float(f"{x:.6g}"))(B)
    assert np.array_equal(printed, DYNREG_B)


def _spd_cases(ct):
    pr = ct.problems
    yield "lap2d_12_nd", pr.laplacian2d_dirichlet(12, 12), ct.grid_nd(12, 12, 1)
    yield "lap3d_10_mmd", pr.laplacian3d_7pt(10, 10, 10, seed=4), "mmd"
    yield "st27_9_nd", pr.stencil3d_27pt(9, 9, 9, seed=3), ct.grid_nd(9, 9, 9)
    yield "lap3d_12_nd", pr.laplacian3d_7pt(12, 12, 12, seed=2), ct.grid_nd(12, 12, 12)


@pytest.mark.parametrize("name", ["lap2d_12_nd", "lap3d_10_mmd", "st27_9_nd", "lap3d_12_nd"])
def test_factor_vs_dense_lapack(ct, name):
    """Device Cholesky and LDL' factors against numpy.linalg.cholesky(PAP')
    mapped into the Dval/Lval layout (unique factors; 1e-12 normwise), and the
    device solve against a dense solve."""
    A, order = {k: (a, o) for k, a, o in _spd_cases(ct)}[name]
    print(f"case {name} n={A.shape[0]}", flush=True)
    F = ct.ChordalCholesky(A, alg=order)
    ct.cholesky_(F)
    M = permuted_dense(A, F.P)
    Lc = np.linalg.cholesky(M)
    D1, L1 = to_layout(F.S, Lc)
    assert rel(F.Dval, D1) <= 1e-12 and rel(F.Lval, L1) <= 1e-12
    B = np.random.default_rng(2).standard_normal((A.shape[0], 3))
    assert rel(ct.solve(F, B), np.linalg.solve(A.toarray(), B)) <= 1e-11
    G = ct.ChordalLDLt(A, F.P, F.S)
    ct.ldlt_(G)
    piv = np.diag(Lc) ** 2
    assert rel(G.d, piv) <= 1e-12
    Lu = Lc / np.diag(Lc)[None, :]
    assert rel(G.Lval, to_layout(F.S, Lu)[1]) <= 1e-12


def _indefinite_cases(ct):
    pr = ct.problems
    yield pr.laplacian2d_dirichlet(10, 10) - 3.5 * sp.identity(100), ct.grid_nd(10, 10, 1)
    yield pr.laplacian3d_7pt(5, 5, 5, seed=2) - 0.4 * sp.identity(125), ct.grid_nd(5, 5, 5)
    yield pr.laplacian3d_7pt(6, 5, 4, seed=8) - 1.5 * sp.identity(120), "mmd"


def test_failure_info_leading_minor(ct):
    """Cholesky of an indefinite matrix on the device: the failing column is
    the first leading principal minor of PAP' that is not positive definite;
    LDL' with signs = +1 fails at the first non-positive pivot (minor ratios)."""
    for A, order in _indefinite_cases(ct):
        A = sp.csc_matrix(A)
        F = ct.ChordalCholesky(A, alg=order)
        ct.cholesky_(F, check=False)
        M = permuted_dense(A, F.P)
        k = first_indefinite_minor(M)
        assert k > 0 and F.info == F.perm[k - 1]
        G = ct.ChordalLDLt(A, F.P, F.S)
        ct.ldlt_(G, signs=np.ones(A.shape[0]), check=False)
        piv = leading_pivots(M)
        kk = int(np.nonzero(piv <= 0)[0][0]) + 1
        assert G.info == G.perm[kk - 1]


@pytest.mark.parametrize("diag", [0, 1])
def test_dynamic_regularization_rule_device(ct, oracle, diag):
    """DynamicRegularization (dynamic_regularization.jl:75-90) on the device, for
    Cholesky and LDL': equal to the oracle, the factor is exact for PAP' + E with
    E diagonal, replaced pivots equal delta and were below epsilon before."""
    pr = ct.problems
    delta, eps = _dyn(ct)
    cases = [(weakly_indefinite(pr.laplacian2d_dirichlet(10, 10), [3, 41, 77], 1), ct.grid_nd(10, 10, 1)),
             (weakly_indefinite(pr.laplacian3d_7pt(5, 5, 5, seed=2), [0, 62, 124, 99], 2),
              ct.grid_nd(5, 5, 5))]
    for A, order in cases:
        cls = ct.ChordalCholesky if diag == 0 else ct.ChordalLDLt
        F = cls(A, alg=order)
        signs = np.ones(A.shape[0])
        if diag == 0:
            ct.cholesky_(F, reg=ct.DynamicRegularization())
        else:
            ct.ldlt_(F, signs=signs, reg=ct.DynamicRegularization())
        assert ct.issuccess(F)
        S = F.S
        Dv, Lv = oracle.assemble(S, F.P, A)
        D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv, diag=diag, signs=signs, reg=(delta, eps))
        assert info0 == 0
        Dg, Lg, dg = F.Dval, F.Lval, F.d
        assert rel(Dg, D0) <= 1e-12 and rel(Lg, L0) <= 1e-12
        if diag == 1:
            assert rel(dg, d0) <= 1e-12
        M = permuted_dense(A, F.P)
        if diag == 0:
            Ld = dense_lower(S, Dg, Lg)
            piv, R = np.diag(Ld) ** 2, Ld @ Ld.T
        else:
            Lu = dense_lower(S, Dg, Lg, unit=True)
            piv, R = dg, Lu @ np.diag(dg) @ Lu.T
        E = R - M
        scale = np.abs(M).max() + np.abs(np.diag(E)).max()
        assert np.abs(E - np.diag(np.diag(E))).max() <= 1e-11 * scale
        e = np.diag(E)
        hit = np.abs(e) > 1e-9 * scale
        assert hit.any()
        assert np.allclose(piv[hit], delta, rtol=1e-12, atol=0)
        assert np.all((piv - e)[hit] < eps) and np.all(piv[~hit] >= eps)
        if diag == 1:
            nod = nod_columns(S)
            assert np.array_equal(dense_lower(S, Dg, Lg).diagonal()[nod], np.diag(M)[nod])


# ---- wide fronts: grouped trailing updates ------------------------------------

@pytest.fixture(scope="module")
def wide(ct, oracle):
    """7-point Laplacian on a 30 x 30 x 28 grid with geometric nested dissection:
    top separator 30 x 30 = 900 columns (15 panel blocks of 64)."""
    A = ct.problems.laplacian3d_7pt(30, 30, 28, seed=12)
    P, S = ct.symbolic(A, alg=ct.grid_nd(30, 30, 28))
    assert S.nn().max() >= 12 * 64
    Dv, Lv = oracle.assemble(S, P, A)
    oracle.use_host_blas(True)
    try:
        chol = oracle.chol_impl(S, Dv, Lv, diag=0)
        ldlt = oracle.chol_impl(S, Dv, Lv, diag=1)
    finally:
        oracle.use_host_blas(False)
    return A, P, S, chol, ldlt


def test_wide_front_cholesky_fp64(ct, oracle, wide):
    A, P, S, (D0, L0, d0, info0), _ = wide
    print(f"wide cholesky n={A.shape[0]} max nn={S.nn().max()}", flush=True)
    assert info0 == 0
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    assert rel(F.Dval, D0) <= 1e-12 and rel(F.Lval, L0) <= 1e-12
    B = np.random.default_rng(5).standard_normal((A.shape[0], 16))
    X = ct.solve(F, B)
    assert rel(X, oracle.solve(S, P, D0, L0, d0, B)) <= 1e-11
    assert np.linalg.norm(A @ X - B) / np.linalg.norm(B) <= 1e-12


def test_wide_front_ldlt_fp64(ct, oracle, wide):
    A, P, S, _, (D0, L0, d0, info0) = wide
    print("wide ldlt", flush=True)
    assert info0 == 0
    F = ct.ChordalLDLt(A, P, S)
    ct.ldlt_(F)
    assert rel(F.d, d0) <= 1e-12 and rel(F.Lval, L0) <= 1e-12 and rel(F.Dval, D0) <= 1e-12
    B = np.random.default_rng(6).standard_normal((A.shape[0], 5))
    X = ct.solve(F, B)
    assert rel(X, oracle.solve(S, P, D0, L0, d0, B, diag=1)) <= 1e-11


def test_wide_front_fp32(ct, oracle, wide):
    """FP32 fronts on the wide problem: factor within FP32 rounding of the FP64
    oracle (normwise 1e-5), and FP64 refinement reaches 1e-10."""
    A, P, S, (D0, L0, d0, _), _ = wide
    print("wide fp32", flush=True)
    F = ct.ChordalCholesky(A, P, S, precision=ct.MF_FP32)
    ct.cholesky_(F)
    assert rel(F.Dval, D0) <= 1e-5 and rel(F.Lval, L0) <= 1e-5
    B = np.random.default_rng(7).standard_normal((A.shape[0], 4))
    X = ct.refine(F, B, tol=1e-10)
    assert rel(X, oracle.solve(S, P, D0, L0, d0, B)) <= 1e-9


def test_fused_diagonal_tasks(ct, oracle, wide):
    """The tile schedule with each diagonal-tile task also running the next
    panel solve and look-ahead update (planner tuning fuse = 1) gives the
    oracle's factor on the wide-front problem, Cholesky and LDLt."""
    A, P, S, (D0, L0, d0, _), (D1, L1, d1, _) = wide
    tuning = [8, 16384, 128, 512, 50000, 0, 1]
    F = ct.ChordalCholesky(A, P, S, tuning=tuning)
    ct.cholesky_(F)
    assert F._plan.tuning()["fuse"] == 1
    assert rel(F.Dval, D0) <= 1e-12 and rel(F.Lval, L0) <= 1e-12
    G = ct.ChordalLDLt(A, P, S, tuning=tuning)
    ct.ldlt_(G)
    assert rel(G.d, d1) <= 1e-12 and rel(G.Lval, L1) <= 1e-12


def test_tuned_parameter_sets(ct, oracle, wide):
    """Every parameter set of the autotuner's table (the planner tunings the
    benchmark runs with) gives the oracle's factor on the wide-front problem,
    FP64 Cholesky and LDLt, and FP32 to single precision."""
    import json
    import os
    A, P, S, (D0, L0, d0, _), (D1, L1, d1, _) = wide
    table = os.path.join(os.path.dirname(ct.__file__), "tuning.json")
    with open(table) as f:
        sets = sorted({tuple(v["tuning"]) for v in json.load(f).values()})
    assert sets
    for tuning in sets:
        print(f"tuning {tuning}", flush=True)
        F = ct.ChordalCholesky(A, P, S, tuning=list(tuning))
        ct.cholesky_(F)
        assert rel(F.Dval, D0) <= 1e-12 and rel(F.Lval, L0) <= 1e-12
        G = ct.ChordalLDLt(A, P, S, tuning=list(tuning))
        ct.ldlt_(G)
        assert rel(G.d, d1) <= 1e-12 and rel(G.Lval, L1) <= 1e-12
        H = ct.ChordalCholesky(A, P, S, tuning=list(tuning), precision=ct.MF_FP32)
        ct.cholesky_(H)
        assert rel(H.Dval, D0) <= 1e-5 and rel(H.Lval, L0) <= 1e-5
