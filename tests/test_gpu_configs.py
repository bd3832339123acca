"""The benchmark configurations at their full sizes on the HIP path.

  config 1  2-D 5-point Dirichlet Laplacian 100 x 100, MMD order, one N(0,1)
            right-hand side (seed 1): factor entry by entry against the oracle
  config 3  3-D 27-point stencil 48^3 (seed 3), FP32 fronts + FP64 iterative
            refinement to 1e-10 for 16 right-hand sides, against the oracle's
            FP64 solution; the plain FP64 factor against the oracle's
  config 4  4,096 uneven grid Laplacians n ~ U{50..2000} as one forest, FP64
            and FP32 + refinement: every system's residual, the oracle on a
            sample of systems
  config 5  KKT LDLt [H A'; A -1e-6 I] at n_H = 1e3, 1e4 and 1e5 (log-spaced
            over the range), one value update each, residual after
            refinement, the oracle at the two smaller sizes
(config 2, the 64^3 Laplacian, is test_repeat_factorization_bitwise.)
"""
import numpy as np
import pytest

from densehelp import rel

pytestmark = pytest.mark.gpu


def test_config1_lap2d_100_mmd(ct, oracle):
    A = ct.problems.laplacian2d_dirichlet(100, 100)
    P, S = ct.symbolic(A, alg="mmd")
    print(f"config 1: n={A.shape[0]} fronts={S.nfr}", flush=True)
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv)
    assert info0 == 0
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    D1, L1 = F.Dval, F.Lval
    assert rel(D1, D0) <= 1e-12 and rel(L1, L0) <= 1e-12
    # entry by entry, relative to the size of the column's diagonal entry
    assert np.abs(D1 - D0).max() <= 1e-13 * np.abs(D0).max()
    assert np.abs(L1 - L0).max() <= 1e-13 * np.abs(D0).max()
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x = ct.solve(F, b)
    assert rel(x, oracle.solve(S, P, D0, L0, d0, b)) <= 1e-12
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-14


@pytest.fixture(scope="module")
def st27(ct, oracle):
    A = ct.problems.stencil3d_27pt(48, 48, 48, seed=3)
    P, S = ct.symbolic(A, alg=ct.grid_nd(48, 48, 48))
    Dv, Lv = oracle.assemble(S, P, A)
    oracle.use_host_blas(True)
    try:
        D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv)
        B = np.random.default_rng(3).standard_normal((A.shape[0], 16))
        X0 = oracle.solve(S, P, D0, L0, d0, B)
    finally:
        oracle.use_host_blas(False)
    assert info0 == 0
    return A, P, S, D0, L0, B, X0


def test_config3_st27_48_mixed(ct, st27):
    A, P, S, D0, L0, B, X0 = st27
    print(f"config 3 mixed: n={A.shape[0]}", flush=True)
    F = ct.ChordalCholesky(A, P, S, precision=ct.MF_FP32)
    ct.cholesky_(F)
    X = ct.refine(F, B, tol=1e-10)
    res = np.linalg.norm(A @ X - B, axis=0) / np.linalg.norm(B, axis=0)
    assert res.max() <= 1e-10
    err = rel(X, X0)
    print(f"config 3: {F.ir_iters} corrections, worst residual {res.max():.2e}, error vs FP64 oracle "
          f"{err:.2e}", flush=True)
    assert err <= 1e-9


def test_config3_st27_48_fp64(ct, st27):
    A, P, S, D0, L0, B, X0 = st27
    print("config 3 fp64", flush=True)
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    assert rel(F.Dval, D0) <= 1e-12 and rel(F.Lval, L0) <= 1e-12
    X = ct.solve(F, B)
    assert rel(X, X0) <= 1e-11


@pytest.fixture(scope="module")
def batch(ct):
    specs = ct.problems.batch_specs(4096, 4)
    mats = [ct.problems.grid_laplacian_random(d, 4000 + k) for k, d in enumerate(specs)]
    orders = [ct.grid_nd(*(d if len(d) == 3 else (d[0], d[1], 1))) for d in specs]
    return mats, orders


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config4_batch4096(ct, oracle, batch, precision):
    mats, orders = batch
    print(f"config 4 {precision}: {len(mats)} systems", flush=True)
    prec = ct.MF_FP64 if precision == "fp64" else ct.MF_FP32
    Bt = ct.BatchedFactorization(mats, orders, precision=prec).factorize()
    rng = np.random.default_rng(4)
    bs = [rng.standard_normal(m.shape[0]) for m in mats]
    xs = Bt.solve(bs) if prec == ct.MF_FP64 else Bt.refine(bs, tol=1e-10)
    res = np.array([np.linalg.norm(m @ x - b) / np.linalg.norm(b) for m, b, x in zip(mats, bs, xs)])
    assert res.max() <= (1e-12 if prec == ct.MF_FP64 else 1e-10), res.max()
    for k in range(0, len(mats), 128):
        A, o, b, x = mats[k], orders[k], bs[k], xs[k]
        P, S = ct.symbolic(A, alg=o)
        Dv, Lv = oracle.assemble(S, P, A)
        D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv)
        assert info0 == 0
        assert rel(x, oracle.solve(S, P, D0, L0, d0, b)) <= (1e-11 if prec == ct.MF_FP64 else 1e-9)


@pytest.mark.parametrize("nh", [1000, 10000, 100000])
def test_config5_kkt_sizes(ct, oracle, nh):
    K0, nh_, m = ct.problems.kkt_matrix(ct.problems.near_cubic_dims(nh, 3), seed=77)
    K = ct.problems.kkt_value_update(K0, nh_, seed=78)
    perm = ct.mmd(K)
    P, S = ct.symbolic(K, alg=perm)
    nj = S.nn() + S.na()
    print(f"config 5: n_H={nh_} n={K.shape[0]} max front {nj.max()}", flush=True)
    F = ct.ChordalLDLt(K, P, S)
    ct.ldlt_(F)
    b = np.random.default_rng(nh).standard_normal(K.shape[0])
    x = ct.refine(F, b, tol=1e-10)
    assert np.linalg.norm(K @ x - b) / np.linalg.norm(b) <= 1e-10
    if nh <= 10000:
        Dv, Lv = oracle.assemble(S, P, K)
        oracle.use_host_blas(True)
        try:
            D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv, diag=1)
        finally:
            oracle.use_host_blas(False)
        assert info0 == 0
        assert rel(F.d, d0) <= 1e-10 and rel(F.Lval, L0) <= 1e-10
        x0 = oracle.solve(S, P, D0, L0, d0, b, diag=1)
        assert rel(x, x0) <= 1e-8
