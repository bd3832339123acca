"""The C-ABI boundary beyond the :L layout: UPLO = :U storage
(chordal_triangular.jl:895-926, 1015-1055; the reference tests both layouts,
test/multifrontal.jl:6), right division B / F (divide.jl:71-80,
test/multifrontal.jl:25), the drop-in entry points called repeatedly on one
plan, and the assembled values read back before a factorization."""
import numpy as np
import pytest
import scipy.sparse as sp

from densehelp import rel, to_upper_layout

pytestmark = pytest.mark.gpu


def _cases(ct):
    pr = ct.problems
    yield "lap3d_12", pr.laplacian3d_7pt(12, 12, 12, seed=2), ct.grid_nd(12, 12, 12)
    yield "st27_8", pr.stencil3d_27pt(8, 8, 8, seed=3), ct.grid_nd(8, 8, 8)
    yield "lap2d_30", pr.laplacian2d_dirichlet(30, 30), "mmd"


@pytest.fixture(scope="module")
def cases(ct):
    return {k: (a, o) for k, a, o in _cases(ct)}


@pytest.mark.parametrize("name", ["lap3d_12", "st27_8", "lap2d_30"])
@pytest.mark.parametrize("diag", [0, 1])
def test_upper_layout_factor(ct, oracle, cases, name, diag):
    """ChordalCholesky{:U} / ChordalLDLt{:U}: Dval/Lval hold the transposed
    blocks of the :L factor (U = L'), solves are those of the :L factor."""
    A, order = cases[name]
    print(f"case {name} diag={diag} uplo=U", flush=True)
    P, S = ct.symbolic(A, alg=order)
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv, diag=diag)
    assert info0 == 0
    DU0, LU0 = to_upper_layout(S, D0, L0)
    cls = ct.ChordalCholesky if diag == 0 else ct.ChordalLDLt
    F = cls(A, P, S, uplo="U")
    # before the factorization the arrays hold the assembled A (copyto!)
    DUa, LUa = to_upper_layout(S, Dv, Lv)
    assert np.array_equal(F.Dval, DUa) and np.array_equal(F.Lval, LUa)
    (ct.cholesky_ if diag == 0 else ct.ldlt_)(F)
    assert rel(F.Dval, DU0) <= 1e-12 and rel(F.Lval, LU0) <= 1e-12
    if diag == 1:
        assert rel(F.d, d0) <= 1e-12
    B = np.random.default_rng(4).standard_normal((A.shape[0], 3))
    X = ct.solve(F, B)
    assert rel(X, oracle.solve(S, P, D0, L0, d0, B, diag=diag)) <= 1e-11
    # the flat-array seam in the :U layout
    D1, L1, d1, info = ct.chol_impl(S, DUa, LUa, diag=diag, uplo="U")
    assert info == 0 and rel(D1, DU0) <= 1e-12 and rel(L1, LU0) <= 1e-12
    C = P.apply(B)
    Y = ct.div_impl(S, DU0, LU0, d0 if diag == 1 else None, C, diag=diag, uplo="U")
    assert rel(Y, oracle.div_natural(S, D0, L0, d0, C, diag=diag)) <= 1e-11


def test_upper_layout_file_round_trip(ct, cases, tmp_path):
    A, order = cases["lap3d_12"]
    F = ct.ChordalLDLt(A, alg=order, uplo="U")
    ct.ldlt_(F)
    ct.save_factor(F, tmp_path / "u.npz")
    G = ct.load_factor(tmp_path / "u.npz")
    assert G.uplo == "U" and np.array_equal(G.Lval, F.Lval) and np.array_equal(G.d, F.d)
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    assert rel(ct.solve(G, b), ct.solve(F, b)) <= 1e-12


@pytest.mark.parametrize("name", ["lap3d_12", "lap2d_30"])
def test_right_division(ct, cases, name):
    """B / F = B A^-1 for a row block B (test/multifrontal.jl:25 pattern:
    (C / F) * M == C), for both layouts and both factorizations."""
    A, order = cases[name]
    C = np.random.default_rng(5).standard_normal((5, A.shape[0]))
    for cls, fac in ((ct.ChordalCholesky, ct.cholesky_), (ct.ChordalLDLt, ct.ldlt_)):
        for uplo in ("L", "U"):
            F = cls(A, alg=order, uplo=uplo)
            fac(F)
            Y = ct.rsolve(C, F)
            assert Y.shape == C.shape
            assert np.linalg.norm(Y @ A - C) / np.linalg.norm(C) <= 1e-12
            assert rel(Y, ct.solve(F, C.T).T) <= 1e-15
            y = ct.rsolve(C[0], F)
            assert rel(y, Y[0]) <= 1e-15


def test_dropin_repeated_calls(ct, oracle, cases):
    """mf_chol_impl / mf_div_impl keep one factor object per plan: calls in a
    row, switching layout and factorization, give the oracle's results."""
    A, order = cases["st27_8"]
    P, S = ct.symbolic(A, alg=order)
    Dv, Lv = oracle.assemble(S, P, A)
    ref = {dg: oracle.chol_impl(S, Dv, Lv, diag=dg) for dg in (0, 1)}
    for dg, uplo in [(0, "L"), (1, "L"), (0, "U"), (1, "U"), (0, "L")]:
        Din, Lin = (Dv, Lv) if uplo == "L" else to_upper_layout(S, Dv, Lv)
        D1, L1, d1, info = ct.chol_impl(S, Din, Lin, diag=dg, uplo=uplo)
        D0, L0, d0, _ = ref[dg]
        if uplo == "U":
            D0, L0 = to_upper_layout(S, D0, L0)
        assert info == 0 and rel(D1, D0) <= 1e-12 and rel(L1, L0) <= 1e-12
        if dg == 1:
            assert rel(d1, ref[1][2]) <= 1e-12
        else:
            assert np.array_equal(d1, np.ones(S.n))


def test_assembled_values_and_refresh(ct, oracle, cases):
    """copyto!(F, A) then reading Dval/Lval gives the assembled PAP' (the
    arrays are produced from the device panels on demand); a refresh with new
    values after a factorization does the same, and the LDLt factor keeps the
    assembled entry on the diagonal of nn = 1 fronts."""
    A, order = cases["lap2d_30"]
    P, S = ct.symbolic(A, alg=order)
    F = ct.ChordalLDLt(A, P, S)
    Dv, Lv = oracle.assemble(S, P, A)
    assert np.array_equal(F.Dval, Dv) and np.array_equal(F.Lval, Lv)
    ct.ldlt_(F)
    A2 = (A + sp.diags(np.linspace(0.5, 1.5, A.shape[0]))).tocsc()
    F.copyto_(A2)
    Dv2, Lv2 = oracle.assemble(S, P, A2)
    assert np.array_equal(F.Dval, Dv2) and np.array_equal(F.Lval, Lv2)
    ct.ldlt_(F)
    D0, L0, d0, _ = oracle.chol_impl(S, Dv2, Lv2, diag=1)
    assert rel(F.Dval, D0) <= 1e-12 and rel(F.Lval, L0) <= 1e-12 and rel(F.d, d0) <= 1e-12
