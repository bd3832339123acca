"""Parity of the HIP engine against the CPU oracle (same perm, same supernode
partition, same assembled values).

Factor entries are compared with the reference's own normwise measure
(isapprox on ChordalTriangular: sqrt(dist2) <= rtol * max(norm),
src/Multifrontal.jl/src/chordal_triangular.jl:648-660,710-713) over Dval and
Lval; FP64 tolerance 1e-12.  Solves are checked against the oracle's solve
and by the residual ||Ax-b||/||b|| (test/multifrontal.jl:23-40 pattern).
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.linalg.norm(a), np.linalg.norm(b), 1e-300)
    return np.linalg.norm(a - b) / den


def diag_mask(S):
    """Mask that hides the diagonal entries of the Dval blocks."""
    m = np.ones(S.ndz, dtype=bool)
    nn = S.nn()
    for j in range(S.nfr):
        p = S.Dptr[j] - 1
        k = nn[j]
        m[p + np.arange(k) * (k + 1)] = False
    return m


def problems(ct):
    P = ct.problems
    out = []
    A5 = sp.csc_matrix(np.array([[4, 2, 0, 0, 2], [2, 5, 0, 0, 3], [0, 0, 4, 2, 0],
                                 [0, 0, 2, 5, 2], [2, 3, 0, 2, 7]], float))
    out.append(("readme5", A5, [3, 1, 2, 4, 5], "maximal"))
    out.append(("lap2d_30_nd", P.laplacian2d_dirichlet(30, 30), ct.grid_nd(30, 30, 1), "maximal"))
    out.append(("lap2d_20_nat", P.laplacian2d_dirichlet(20, 20), None, "maximal"))
    out.append(("lap2d_12_nodal", P.laplacian2d_dirichlet(12, 12), ct.grid_nd(12, 12, 1), "nodal"))
    out.append(("lap3d_12_fund", P.laplacian3d_7pt(12, 12, 12, seed=2), ct.grid_nd(12, 12, 12),
                "fundamental"))
    out.append(("lap3d_16", P.laplacian3d_7pt(16, 16, 16, seed=2), ct.grid_nd(16, 16, 16), "maximal"))
    out.append(("st27_10", P.stencil3d_27pt(10, 10, 10, seed=3), ct.grid_nd(10, 10, 10), "maximal"))
    out.append(("lap3d_20x18x9", P.laplacian3d_7pt(20, 18, 9, seed=5), ct.grid_nd(20, 18, 9), "maximal"))
    return out


PROBS = ["readme5", "lap2d_30_nd", "lap2d_20_nat", "lap2d_12_nodal", "lap3d_12_fund", "lap3d_16",
         "st27_10", "lap3d_20x18x9"]


@pytest.fixture(scope="module")
def cases(ct):
    return {name: (A, order, snd) for name, A, order, snd in problems(ct)}


@pytest.mark.parametrize("name", PROBS)
def test_cholesky_factor_parity(ct, oracle, cases, name):
    A, order, snd = cases[name]
    print(f"case {name} n={A.shape[0]}", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv, diag=0)
    assert info0 == 0
    # drop-in seam (chol_impl! arguments)
    D1, L1, d1, info1 = ct.chol_impl(S, Dv, Lv, diag=ct.MF_CHOLESKY)
    assert info1 == 0
    assert rel(D1, D0) <= RTOL64 and rel(L1, L0) <= RTOL64
    # API path: assembly on the device from flatindices
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    assert ct.issuccess(F)
    assert rel(F.Dval, D0) <= RTOL64 and rel(F.Lval, L0) <= RTOL64
    # the upper triangles of the diagonal blocks stay zero
    nn = S.nn()
    for j in range(0, S.nfr, max(1, S.nfr // 50)):
        k = nn[j]
        blk = F.Dval[S.Dptr[j] - 1: S.Dptr[j] - 1 + k * k].reshape(k, k, order="F")
        assert np.all(np.triu(blk, 1) == 0)


@pytest.mark.parametrize("name", PROBS)
@pytest.mark.parametrize("nrhs", [1, 16, 21])
def test_cholesky_solve_parity(ct, oracle, cases, name, nrhs):
    A, order, snd = cases[name]
    print(f"case {name} nrhs={nrhs}", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    rng = np.random.default_rng(nrhs)
    B = rng.standard_normal((A.shape[0], nrhs))
    X = ct.solve(F, B)
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, _ = oracle.chol_impl(S, Dv, Lv, diag=0)
    X0 = oracle.solve(S, P, D0, L0, d0, B, diag=0)
    assert rel(X, X0) <= 1e-11
    res = np.linalg.norm(A @ X - B) / np.linalg.norm(B)
    assert res <= 1e-12
    # natural solve through the div_impl! seam
    C = P.apply(B)
    Xn = ct.div_impl(S, D0, L0, None, C, diag=ct.MF_CHOLESKY)
    Xn0 = oracle.div_natural(S, D0, L0, d0, C, diag=0)
    assert rel(Xn, Xn0) <= 1e-11


@pytest.mark.parametrize("name", ["lap2d_30_nd", "lap3d_16", "st27_10", "lap3d_12_fund"])
def test_ldlt_parity(ct, oracle, cases, name):
    A, order, snd = cases[name]
    print(f"case {name} ldlt", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv, diag=1)
    assert info0 == 0
    F = ct.ChordalLDLt(A, P, S)
    ct.ldlt_(F)
    assert ct.issuccess(F)
    assert rel(F.d, d0) <= RTOL64
    assert rel(F.Lval, L0) <= RTOL64
    # Dval: off-diagonal entries pinned; diagonal entries hold the pivot before
    # regularization (nn > 1) or the assembled entry of A (nn = 1)
    assert rel(F.Dval, D0) <= RTOL64
    B = np.random.default_rng(1).standard_normal((A.shape[0], 3))
    X = ct.solve(F, B)
    assert np.linalg.norm(A @ X - B) / np.linalg.norm(B) <= 1e-12


def test_ldlt_quasidefinite_kkt(ct, oracle):
    K, nh, m = ct.problems.kkt_matrix((6, 6, 5), seed=5)
    P, S = ct.symbolic(K, alg="mmd") if _has_mmd(ct) else ct.symbolic(K, alg=_kkt_order(nh, m))
    Dv, Lv = oracle.assemble(S, P, K)
    D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv, diag=1)
    assert info0 == 0
    F = ct.ChordalLDLt(K, P, S)
    signs = np.concatenate([np.ones(nh), -np.ones(m)])
    ct.ldlt_(F, signs=signs)
    assert rel(F.d, d0) <= RTOL64 and rel(F.Lval, L0) <= RTOL64
    assert np.all(np.sign(F.d[P.invp - 1]) == signs)
    b = np.random.default_rng(2).standard_normal(K.shape[0])
    x = ct.solve(F, b)
    # the pivots of this ordering span 1e-6 .. 4e6, so the attainable residual
    # is the oracle's own (about 2e-10), not machine precision
    x0 = oracle.solve(S, P, D0, L0, d0, b, diag=1)
    res0 = np.linalg.norm(K @ x0 - b) / np.linalg.norm(b)
    assert rel(x, x0) <= 1e-9
    assert np.linalg.norm(K @ x - b) / np.linalg.norm(b) <= max(1e-12, 4 * res0)


def _has_mmd(ct):
    try:
        ct.mmd(np.eye(2))
        return True
    except NotImplementedError:
        return False


def _kkt_order(nh, m):
    # H block first, constraint rows last: quasi-definite, so any order works
    return np.arange(1, nh + m + 1)


def test_readme_examples(ct):
    """Known answers printed in src/Multifrontal.jl/README.md:29-124."""
    A = np.array([[4, 2, 0, 0, 2], [2, 5, 0, 0, 3], [0, 0, 4, 2, 0], [0, 0, 2, 5, 2],
                  [2, 3, 0, 2, 7]], float)
    F = ct.ChordalCholesky(A)
    ct.cholesky_(F)
    assert np.allclose(ct.solve(F, np.array([4., 2, 0, 0, 2])), [1, 0, 0, 0, 0], atol=1e-14)
    Ai = np.array([[-4, -2, 0, 0, 2], [-2, -5, 0, 0, 3], [0, 0, 4, 2, 0], [0, 0, 2, 5, 2],
                   [2, 3, 0, 2, 7]], float)
    G = ct.ChordalLDLt(Ai, alg=[3, 1, 2, 4, 5])
    ct.ldlt_(G)
    assert np.allclose(ct.solve(G, np.array([-4., -2, 0, 0, 2])), [1, 0, 0, 0, 0], atol=1e-14)
    assert sorted(G.d.tolist()) == sorted([4.0, -4.0, -4.0, 4.0, 8.0])
    G2 = ct.ChordalLDLt(Ai, alg=[3, 1, 2, 4, 5])
    ct.ldlt_(G2, signs=[-1, -1, 1, 1, 1])
    assert ct.issuccess(G2)


@pytest.mark.parametrize("name", ["lap2d_30_nd", "lap3d_16"])
def test_failure_info(ct, oracle, cases, name):
    """Not positive definite: info is the first failing column (cholesky.jl:384-388),
    mapped through perm (cholesky.jl:250-254); check=True raises PosDefException."""
    A, order, snd = cases[name]
    lam = 4.0 if name.startswith("lap2d") else 0.5
    B = (A - lam * sp.identity(A.shape[0])).tocsc()
    P, S = ct.symbolic(B, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, B)
    _, _, _, info0 = oracle.chol_impl(S, Dv, Lv, diag=0)
    assert info0 > 0
    F = ct.ChordalCholesky(B, P, S)
    ct.cholesky_(F, check=False)
    assert not ct.issuccess(F)
    assert F.info == P.perm[info0 - 1]
    F2 = ct.ChordalCholesky(B, P, S)
    with pytest.raises(ct.PosDefException) as e:
        ct.cholesky_(F2)
    assert e.value.info == info0


def test_pattern_reuse(ct, oracle):
    """flatindices -> fill!(F, 0) -> factor fails -> new values -> succeeds
    (test/multifrontal.jl:76-113)."""
    A = ct.problems.laplacian3d_7pt(10, 10, 10, seed=7)
    P, S = ct.symbolic(A, alg=ct.grid_nd(10, 10, 10))
    F = ct.ChordalCholesky(A * 0.0, P, S)
    ct.cholesky_(F, check=False)
    assert not ct.issuccess(F)
    for k in range(3):
        Ak = ct.problems.laplacian3d_7pt(10, 10, 10, seed=10 + k)
        F.copyto_(Ak)
        ct.cholesky_(F)
        Dv, Lv = oracle.assemble(S, P, Ak)
        D0, L0, _, _ = oracle.chol_impl(S, Dv, Lv)
        assert rel(F.Lval, L0) <= RTOL64 and rel(F.Dval, D0) <= RTOL64


def test_dynamic_regularization(ct, oracle):
    """DynamicRegularization example (dynamic_regularization.jl:12-30)."""
    A = np.array([[1, 1, 2], [1, 1, 3], [2, 3, 1]], float)
    P, S = ct.symbolic(A)
    Dv, Lv = oracle.assemble(S, P, A)
    reg = ct.DynamicRegularization()
    delta, eps = reg.initialize(np.finfo(float).eps)
    D0, L0, _, info0 = oracle.chol_impl(S, Dv, Lv, diag=0, signs=np.ones(3), reg=(delta, eps))
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F, reg=reg)
    assert ct.issuccess(F) and info0 == 0
    assert rel(F.Dval, D0) <= RTOL64 and rel(F.Lval, L0) <= RTOL64


@pytest.mark.parametrize("name", ["lap3d_16", "st27_10"])
def test_fp32_fronts(ct, oracle, cases, name):
    """FP32 fronts: factor within FP32 rounding of the FP64 oracle."""
    A, order, snd = cases[name]
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, _, _ = oracle.chol_impl(S, Dv, Lv)
    F = ct.ChordalCholesky(A, P, S, precision=ct.MF_FP32)
    ct.cholesky_(F)
    assert rel(F.Lval, L0) <= 1e-4 and rel(F.Dval, D0) <= 1e-4
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    x = ct.solve(F, b)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-3


@pytest.mark.parametrize("name", ["lap3d_16", "st27_10", "lap3d_20x18x9"])
def test_mixed_precision_refinement(ct, oracle, cases, name):
    """FP32 fronts + FP64 iterative refinement reach 1e-10 relative residual
    for 16 right-hand sides and agree with the FP64 oracle's solution; the
    value refresh also updates the refinement matrix."""
    A, order, snd = cases[name]
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, _ = oracle.chol_impl(S, Dv, Lv)
    B = np.random.default_rng(5).standard_normal((A.shape[0], 16))
    X0 = oracle.solve(S, P, D0, L0, d0, B)
    F = ct.ChordalCholesky(A, P, S, precision=ct.MF_FP32)
    ct.cholesky_(F)
    X = ct.refine(F, B, tol=1e-10)
    res = np.linalg.norm(A @ X - B, axis=0) / np.linalg.norm(B, axis=0)
    assert res.max() <= 1e-10 and 1 <= F.ir_iters <= 10
    assert rel(X, X0) <= 1e-9
    # new values on the same pattern: refresh, refactor, refine again
    A2 = (A + sp.diags(np.full(A.shape[0], 0.5))).tocsc()
    F.copyto_(A2)
    ct.cholesky_(F)
    X2 = ct.refine(F, B[:, :3], tol=1e-10)
    res2 = np.linalg.norm(A2 @ X2 - B[:, :3], axis=0) / np.linalg.norm(B[:, :3], axis=0)
    assert res2.max() <= 1e-10


@pytest.mark.parametrize("name", ["lap3d_16", "st27_10"])
def test_fp32_factor_fp64_solves(ct, oracle, cases, name):
    """FP32 fronts with FP64-arithmetic solves on the widened factor: the
    refinement reaches 1e-10 in no more corrections than with FP32 solves,
    agrees with the FP64 oracle, and a new factorization refreshes the copy."""
    A, order, snd = cases[name]
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, _ = oracle.chol_impl(S, Dv, Lv)
    B = np.random.default_rng(6).standard_normal((A.shape[0], 16))
    X0 = oracle.solve(S, P, D0, L0, d0, B)
    F32 = ct.ChordalCholesky(A, P, S, precision=ct.MF_FP32)
    ct.cholesky_(F32)
    ct.refine(F32, B, tol=1e-10)
    F = ct.ChordalCholesky(A, P, S, precision=ct.MF_FP32, solve_fp64=True)
    ct.cholesky_(F)
    # the FP32 factor itself is unchanged by the FP64 solves
    assert np.array_equal(F.Dval, F32.Dval) and np.array_equal(F.Lval, F32.Lval)
    X = ct.refine(F, B, tol=1e-10)
    res = np.linalg.norm(A @ X - B, axis=0) / np.linalg.norm(B, axis=0)
    assert res.max() <= 1e-10 and F.ir_iters <= F32.ir_iters
    assert rel(X, X0) <= 1e-9
    # one solve runs the FP64 kernels (a result no FP32 solve gives) and is
    # as close to the FP64 solution as one with FP32 arithmetic, or closer
    Y64, Y32 = ct.solve(F, B), ct.solve(F32, B)
    assert not np.array_equal(Y64, Y32)
    assert not np.array_equal(Y64, Y64.astype(np.float32).astype(np.float64))
    assert rel(Y64, X0) <= 1.5 * rel(Y32, X0)
    A2 = (A + sp.diags(np.full(A.shape[0], 0.5))).tocsc()
    F.copyto_(A2)
    ct.cholesky_(F)
    X2 = ct.refine(F, B[:, :3], tol=1e-10)
    res2 = np.linalg.norm(A2 @ X2 - B[:, :3], axis=0) / np.linalg.norm(B[:, :3], axis=0)
    assert res2.max() <= 1e-10


def test_refinement_fp64_and_vector(ct, cases):
    """With FP64 fronts one solve already meets the tolerance (0 corrections);
    a vector right-hand side keeps its shape."""
    A, order, snd = cases["lap3d_16"]
    P, S = ct.symbolic(A, alg=order, snd=snd)
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    b = np.random.default_rng(9).standard_normal(A.shape[0])
    x = ct.refine(F, b, tol=1e-10)
    assert x.shape == b.shape and F.ir_iters == 0
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-12


def _batch(ct, count, seed, nmax=400):
    specs = ct.problems.batch_specs(count, seed, nmin=20, nmax=nmax)
    mats = [ct.problems.grid_laplacian_random(d, seed + k) for k, d in enumerate(specs)]
    orders = [ct.grid_nd(*(d if len(d) == 3 else (d[0], d[1], 1))) for d in specs]
    return mats, orders


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_forest(ct, oracle, precision):
    """A batch of uneven 2-D/3-D systems factored as one forest: every
    instance matches its own oracle factorization and solve."""
    mats, orders = _batch(ct, 24, 40)
    prec = ct.MF_FP64 if precision == "fp64" else ct.MF_FP32
    Bt = ct.BatchedFactorization(mats, orders, precision=prec).factorize()
    rng = np.random.default_rng(1)
    bs = [rng.standard_normal(m.shape[0]) for m in mats]
    xs = Bt.solve(bs) if prec == ct.MF_FP64 else Bt.refine(bs, tol=1e-10)
    for A, o, b, x in zip(mats, orders, bs, xs):
        P, S = ct.symbolic(A, alg=o)
        Dv, Lv = oracle.assemble(S, P, A)
        D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv)
        x0 = oracle.solve(S, P, D0, L0, d0, b)
        assert info0 == 0
        assert rel(x, x0) <= (1e-11 if prec == ct.MF_FP64 else 1e-9)
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-10


def test_batched_failure_names_instance(ct):
    mats, orders = _batch(ct, 6, 41, nmax=120)
    bad = mats[3].tolil()
    bad[5, 5] = -50.0
    mats[3] = bad.tocsc()
    with pytest.raises(ct.BatchFailure) as ei:
        ct.BatchedFactorization(mats, orders).factorize()
    assert ei.value.instance == 3


def test_serving_loop_same_alone_and_in_the_mix(ct, oracle):
    """Continuous batching of KKT LDLt requests (3 patterns x 6 value updates,
    Poisson arrivals): every result equals the request served alone (one at
    a time, no batching), the refined solves meet 1e-10 in relative residual
    and agree with the oracle's LDLt solve (same MMD order) and a dense solve,
    the device-reported residuals agree with the host's, and one correction
    improves on the plain drop-in solve."""
    srv = ct.serving
    reqs = []
    rng = np.random.default_rng(11)
    for p, nh in enumerate([300, 700, 1500]):
        K0, nh_, m = ct.problems.kkt_matrix(ct.problems.near_cubic_dims(nh, 3), seed=20 + p)
        for u in range(6):
            K = ct.problems.kkt_value_update(K0, nh_, seed=100 * p + u)
            reqs.append(srv.Request(K=K, B=rng.standard_normal((K.shape[0], 1)), tag=(p, u)))
    t = srv.poisson_arrivals(len(reqs), rate=2000.0, seed=3)
    order = rng.permutation(len(reqs))
    for r, k in zip(reqs, order):
        r.arrival = float(t[k])
    loop = srv.ServingLoop(diag=ct.MF_LDLT, max_batch=4, streams=3, hbm_budget=2 << 30)
    out = loop.run(reqs)
    assert max(r.batch for r in out) > 1          # some requests were batched
    alone = loop.run([srv.Request(K=r.K, B=r.B, arrival=0.0, tag=r.tag) for r in out],
                     realtime=False, batching=False)
    by_tag = {r.tag: r for r in alone}
    for r in out:
        assert r.info == 0
        a = by_tag[r.tag]
        assert a.batch == 1 and rel(r.X, a.X) <= 1e-12
        res = np.linalg.norm(r.K @ r.X - r.B) / np.linalg.norm(r.B)
        assert res <= 1e-10, (r.tag, res)
        # equal up to the rounding of the two residual sums (orders differ)
        assert abs(r.relres - res) <= 1e-2 * res + 1e-13, (r.relres, res)
        perm = ct.mmd(r.K)
        P, S = ct.symbolic(r.K, alg=perm)
        F = ct.ChordalLDLt(r.K, P, S)
        ct.ldlt_(F)
        x0 = ct.solve(F, r.B)
        res0 = np.linalg.norm(r.K @ x0 - r.B) / np.linalg.norm(r.B)
        assert res <= max(1e-14, res0)
        Dv, Lv = oracle.assemble(S, P, r.K)
        D0, L0, d0, info0 = oracle.chol_impl(S, Dv, Lv, diag=1)
        assert info0 == 0
        xo = oracle.solve(S, P, D0, L0, d0, r.B, diag=1)
        assert rel(r.X, xo) <= 1e-8, (r.tag, rel(r.X, xo))
        assert rel(r.X, np.linalg.solve(r.K.toarray(), r.B)) <= 1e-8
    st = srv.latency_stats(out)
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_edge_cases(ct, oracle):
    # 1x1, diagonal matrix (forest of single-column fronts), disconnected blocks
    for A in [sp.csc_matrix(np.array([[2.0]])), sp.diags(np.arange(1.0, 40.0)).tocsc(),
              sp.block_diag([ct.problems.laplacian2d_dirichlet(5, 5),
                             ct.problems.laplacian3d_7pt(4, 4, 4, seed=1)]).tocsc()]:
        P, S = ct.symbolic(A)
        Dv, Lv = oracle.assemble(S, P, A)
        D0, L0, d0, _ = oracle.chol_impl(S, Dv, Lv)
        F = ct.ChordalCholesky(A, P, S)
        ct.cholesky_(F)
        assert rel(F.Dval, D0) <= RTOL64 and rel(F.Lval, L0) <= RTOL64
        b = np.arange(1.0, A.shape[0] + 1)
        assert rel(ct.solve(F, b), oracle.solve(S, P, D0, L0, d0, b)) <= 1e-12


def test_repeat_factorization_bitwise(ct):
    """The schedule is dynamic but every tile's arithmetic is fixed (update
    groups in order, extend-add sums in child order), so repeated
    factorizations of the 64^3 Laplacian must agree bit for bit.  The loop
    is the device-resident step (value refresh, asynchronous factorization,
    16-RHS device solve, all on one stream without host synchronization);
    every step's factor and residual are checked."""
    import ctypes
    import torch
    from cliquetrees_amd import _native
    A = ct.problems.laplacian3d_7pt(64, 64, 64, seed=2)
    n = A.shape[0]
    P, S = ct.symbolic(A, alg=ct.grid_nd(64, 64, 64))
    F = ct.ChordalCholesky(A, P, S)
    lib = _native.hip()
    h = F._f.handle
    stream = torch.cuda.Stream()
    assert lib.mf_factor_set_stream(h, ctypes.c_void_p(stream.cuda_stream)) == 0
    flat = ct.flatindices(S, A)
    assert lib.mf_factor_set_pattern(h, len(flat), flat.ctypes.data_as(_native.i64p)) == 0
    perm = np.ascontiguousarray(P.perm, dtype=np.int64)
    B0 = np.random.default_rng(4).standard_normal((n, 16))
    with torch.cuda.stream(stream):
        nz = torch.tensor(A.data, dtype=torch.float64, device="cuda")
        Bd0 = torch.tensor(B0.T.copy(), dtype=torch.float64, device="cuda")
        reps = 8
        Ds = [torch.empty(S.ndz, dtype=torch.float64, device="cuda") for _ in range(reps)]
        Ls = [torch.empty(S.nlz, dtype=torch.float64, device="cuda") for _ in range(reps)]
        Xs = [torch.empty_like(Bd0) for _ in range(reps)]
        for rep in range(reps):
            assert lib.mf_factor_refresh(h, ctypes.c_void_p(nz.data_ptr()), 1) == 0
            assert lib.mf_factorize_async(h, ct.MF_CHOLESKY, None, 0, 0.0, 0.0) == 0
            Xs[rep].copy_(Bd0)
            assert lib.mf_solve(h, perm.ctypes.data_as(_native.i64p), ctypes.c_void_p(Xs[rep].data_ptr()),
                                16, n, 1) == 0
            assert lib.mf_factor_get(h, ctypes.c_void_p(Ds[rep].data_ptr()),
                                     ctypes.c_void_p(Ls[rep].data_ptr()), None, 1) == 0
    torch.cuda.synchronize()
    info = ctypes.c_int64(-1)
    assert lib.mf_factor_info(h, ctypes.byref(info)) == 0 and info.value == 0
    for rep in range(reps):
        X = Xs[rep].T.cpu().numpy()
        res = np.linalg.norm(A @ X - B0) / np.linalg.norm(B0)
        same = torch.equal(Ds[rep], Ds[0]) and torch.equal(Ls[rep], Ls[0])
        assert same and res <= 1e-12, (rep, same, res)
        assert torch.equal(Xs[rep], Xs[0]), rep
    # back-to-back solves with the last factor: the persistent solve kernels
    # publish diagonal-block progress across workgroups, and every solve must
    # give the same bits
    with torch.cuda.stream(stream):
        Ys = [torch.empty_like(Bd0) for _ in range(32)]
        for Y in Ys:
            Y.copy_(Bd0)
            assert lib.mf_solve(h, perm.ctypes.data_as(_native.i64p), ctypes.c_void_p(Y.data_ptr()),
                                16, n, 1) == 0
    torch.cuda.synchronize()
    bad = [i for i, Y in enumerate(Ys) if not torch.equal(Y, Xs[0])]
    assert not bad, bad


@pytest.mark.parametrize("name", ["lap2d_30_nd", "lap3d_16"])
def test_factor_file_round_trip(ct, oracle, cases, name, tmp_path):
    """save_factor / load_factor: the restored factor holds the same values bit
    for bit and solves like the factor it was saved from (its diagonal-block
    inverses are recomputed, so to rounding) and like the oracle."""
    A, order, snd = cases[name]
    print(f"case {name}", flush=True)
    P, S = ct.symbolic(A, alg=order, snd=snd)
    F = ct.ChordalCholesky(A, P, S)
    ct.cholesky_(F)
    path = tmp_path / "factor.npz"
    ct.save_factor(F, path)
    G = ct.load_factor(path)
    assert isinstance(G, ct.ChordalCholesky) and G.info == 0
    assert np.array_equal(G.Dval, F.Dval) and np.array_equal(G.Lval, F.Lval)
    assert np.array_equal(G.perm, F.perm)
    B = np.random.default_rng(7).standard_normal((A.shape[0], 16))
    X, Y = ct.solve(F, B), ct.solve(G, B)
    assert rel(Y, X) <= 1e-12
    Dv, Lv = oracle.assemble(S, P, A)
    D0, L0, d0, _ = oracle.chol_impl(S, Dv, Lv, diag=0)
    assert rel(Y, oracle.solve(S, P, D0, L0, d0, B, diag=0)) <= 1e-11
    assert np.linalg.norm(A @ Y - B) / np.linalg.norm(B) <= 1e-12


def test_factor_file_round_trip_ldlt(ct, oracle, tmp_path):
    K, nh, m = ct.problems.kkt_matrix((6, 6, 5), seed=5)
    P, S = ct.symbolic(K, alg="mmd") if _has_mmd(ct) else ct.symbolic(K, alg=_kkt_order(nh, m))
    F = ct.ChordalLDLt(K, P, S)
    ct.ldlt_(F)
    path = tmp_path / "factor_ldlt.npz"
    ct.save_factor(F, path)
    G = ct.load_factor(path)
    assert isinstance(G, ct.ChordalLDLt)
    assert np.array_equal(G.d, F.d) and np.array_equal(G.Lval, F.Lval)
    b = np.random.default_rng(3).standard_normal(K.shape[0])
    x, y = ct.solve(F, b), ct.solve(G, b)
    assert rel(y, x) <= 1e-10
    # a file of another format is refused
    np.savez(tmp_path / "bad.npz", format=np.array("something-else"))
    with pytest.raises(ValueError):
        ct.load_factor(tmp_path / "bad.npz")
