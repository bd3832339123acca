"""Host symbolic analysis (libmfsym.so) against invariants, brute force and
the known answers held in the reference's documentation."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

A5 = np.array([[4, 2, 0, 0, 2], [2, 5, 0, 0, 3], [0, 0, 4, 2, 0], [0, 0, 2, 5, 2],
               [2, 3, 0, 2, 7]], float)
G8 = np.array([[0, 1, 0, 0, 0, 0, 0, 0], [1, 0, 1, 0, 0, 1, 0, 0], [0, 1, 0, 1, 0, 1, 1, 1],
               [0, 0, 1, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 1, 1, 0], [0, 1, 1, 0, 1, 0, 0, 0],
               [0, 0, 1, 0, 1, 0, 0, 1], [0, 0, 1, 0, 0, 0, 1, 0]], float)


def fill_pattern(A, perm):
    """Lower-triangular pattern of the Cholesky factor of PAP' (dense elimination)."""
    n = A.shape[0]
    B = (np.abs(np.asarray(A)[np.ix_(perm - 1, perm - 1)]) > 0) | np.eye(n, dtype=bool)
    L = np.tril(B)
    for k in range(n):
        rows = np.nonzero(L[k + 1:, k])[0] + k + 1
        for i in rows:
            L[rows[rows >= i], i] = True
    return L


def check_invariants(ct, A, P, S):
    n = S.n
    perm = P.perm
    assert sorted(perm.tolist()) == list(range(1, n + 1))
    assert np.all(P.invp[perm - 1] == np.arange(1, n + 1))
    assert S.resptr[0] == 1 and S.resptr[-1] == n + 1 and np.all(np.diff(S.resptr) > 0)
    L = fill_pattern(A.toarray() if sp.issparse(A) else A, perm)
    nn, na = S.nn(), S.na()
    for j in range(1, S.nfr + 1):
        res = S.res(j)
        sep = S.sep(j)
        assert np.all(np.diff(sep) > 0) and (len(sep) == 0 or sep[0] > res[-1])
        # the structure of the first column of the front is res U sep
        col = np.nonzero(L[:, res[0] - 1])[0] + 1
        assert np.array_equal(col, np.concatenate([res, sep]))
        # every column of the supernode has the nested structure
        for t, c in enumerate(res):
            colc = np.nonzero(L[:, c - 1])[0] + 1
            assert np.array_equal(colc, np.concatenate([res[t:], sep]))
        # rel maps sep(child) into bag(j)
        bag = np.concatenate([res, sep])
        for i in S.children(j):
            assert S.pnt[i - 1] == j
            assert np.array_equal(bag[S.rel(i) - 1], S.sep(i))
    assert np.all(np.diff(S.Dptr) == nn * nn) and np.all(np.diff(S.Lptr) == nn * na)
    # children precede parents (postorder) and the root list is the last CSR row
    for j in range(1, S.nfr + 1):
        for i in S.children(j):
            assert i < j
    roots = S.chdtgt[S.chdptr[S.nfr] - 1: S.chdptr[S.nfr + 1] - 1]
    assert np.array_equal(np.sort(roots), np.nonzero(S.pnt == 0)[0] + 1)


def maximal_partition_brute(L):
    """Maximal supernodes by the definition: column v+1 joins v's supernode when
    struct(v) = {v} U struct(v+1) (contiguous columns)."""
    n = L.shape[0]
    starts = [0]
    for v in range(n - 1):
        sv = set(np.nonzero(L[:, v])[0])
        sw = set(np.nonzero(L[:, v + 1])[0])
        if not (sv == sw | {v}):
            starts.append(v + 1)
    return starts


@pytest.mark.parametrize("order", [None, [3, 1, 2, 4, 5], [5, 4, 3, 2, 1]])
def test_readme_5x5_invariants(ct, order):
    P, S = ct.symbolic(A5, alg=order)
    check_invariants(ct, sp.csc_matrix(A5), P, S)


def test_readme_5x5_pattern(ct):
    """symbolic(A) pattern printed in chordal_symbolic.jl:49-57 (10 stored
    entries) for the permutation under which the printed factor satisfies
    P A P' = L L' (src/Multifrontal.jl/README.md:29-44)."""
    P, S = ct.symbolic(A5, alg=[3, 1, 2, 4, 5])
    assert S.ndz + S.nlz >= 10
    L = fill_pattern(A5, P.perm)
    assert L.sum() == 10


def test_clique_tree_docstring_shape(ct):
    """The 8-vertex example of cliquetree (src/clique_trees.jl:47-73): 6
    bags of sizes {3,3,2,3,2,3}; the MMD order of src/MMDLib.jl/README.md:12-32
    (invp = [2,3,6,1,5,7,8,4]) is one order whose clique tree has that shape."""
    invp = np.array([2, 3, 6, 1, 5, 7, 8, 4])
    perm = np.empty(8, dtype=np.int64)
    perm[invp - 1] = np.arange(1, 9)
    P, S = ct.symbolic(G8, alg=perm)
    check_invariants(ct, sp.csc_matrix(G8), P, S)
    bags = sorted((S.nn() + S.na()).tolist())
    assert S.nfr == 6 and bags == sorted([3, 3, 2, 3, 2, 3])


def test_brute_force_small_graphs(ct):
    rng = np.random.default_rng(0)
    for trial in range(40):
        n = int(rng.integers(1, 9))
        M = (rng.random((n, n)) < 0.35).astype(float)
        M = np.triu(M, 1)
        M = M + M.T + np.eye(n)
        order = rng.permutation(n) + 1
        for snd in ["maximal", "nodal", "fundamental"]:
            P, S = ct.symbolic(M, alg=order, snd=snd)
            check_invariants(ct, sp.csc_matrix(M), P, S)
            if snd == "nodal":
                assert S.nfr == n
            if snd == "maximal":
                L = fill_pattern(M, P.perm)
                # the partition is at least as coarse as fundamental and each
                # merged pair satisfies the maximal condition
                starts = S.resptr[:-1] - 1
                for j in range(S.nfr):
                    r0, r1 = S.resptr[j] - 1, S.resptr[j + 1] - 1
                    for v in range(r0, r1 - 1):
                        sv = set(np.nonzero(L[:, v])[0])
                        sw = set(np.nonzero(L[:, v + 1])[0])
                        assert sv == sw | {v}
                assert len(starts) >= 1


def test_exhaustive_orders_4(ct):
    M = np.array([[1, 1, 0, 1], [1, 1, 1, 0], [0, 1, 1, 1], [1, 0, 1, 1]], float)
    for order in itertools.permutations(range(1, 5)):
        P, S = ct.symbolic(M, alg=list(order))
        check_invariants(ct, sp.csc_matrix(M), P, S)


def test_grid_nd_is_permutation(ct):
    for dims in [(1, 1, 1), (7, 1, 1), (5, 6, 1), (4, 5, 6), (16, 16, 16)]:
        o = ct.grid_nd(*dims)
        assert sorted(o.tolist()) == list(range(1, int(np.prod(dims)) + 1))


def test_laplacian_invariants(ct):
    A = ct.problems.laplacian2d_dirichlet(9, 7)
    for order in [None, ct.grid_nd(9, 7, 1)]:
        P, S = ct.symbolic(A, alg=order)
        check_invariants(ct, A, P, S)
    A3 = ct.problems.stencil3d_27pt(4, 5, 3, seed=1)
    P, S = ct.symbolic(A3, alg=ct.grid_nd(4, 5, 3))
    check_invariants(ct, A3, P, S)


def test_flatindices(ct):
    A = ct.problems.laplacian3d_7pt(5, 4, 3, seed=3)
    P, S = ct.symbolic(A, alg=ct.grid_nd(5, 4, 3))
    flat = ct.flatindices(S, A)
    C = A.tocsc()
    C.sort_indices()
    rows = C.indices
    cols = np.repeat(np.arange(C.shape[0]), np.diff(C.indptr))
    v = P.invp[rows] - 1
    w = P.invp[cols] - 1
    assert np.all((flat > 0) == (v >= w))
    vals = np.zeros(S.ndz + S.nlz)
    vals[flat[flat > 0] - 1] = C.data[flat > 0]
    # slots hold exactly the lower triangle of PAP' (compare against the oracle's copyto!)
    import oracle
    Dv, Lv = oracle.assemble(S, P, A)
    assert np.array_equal(vals, np.concatenate([Dv, Lv]))


def test_invalid_input(ct):
    with pytest.raises(ValueError):
        ct.symbolic(A5, alg=[1, 1, 2, 3, 4])
    with pytest.raises(ValueError):
        ct.symbolic(A5, alg=[1, 2, 3])


def test_mmd_readme_golden(ct):
    """src/MMDLib.jl/README.md:12-32: invp = mmd(8, colptr, rowval)."""
    perm = ct.mmd(G8)
    invp = np.empty(8, dtype=np.int64)
    invp[perm - 1] = np.arange(1, 9)
    assert invp.tolist() == [2, 3, 6, 1, 5, 7, 8, 4]


def _fill(A, perm):
    return int(fill_pattern(A, perm).sum())


@pytest.mark.parametrize("delta", [0, 1, 3])
def test_mmd_properties(ct, delta):
    """Valid permutation; isolated vertices first in vertex order; the first
    vertex eliminated has minimum degree; far less fill than the natural order
    on a grid; the symbolic factorization built on it satisfies the invariants."""
    A = ct.problems.laplacian2d_dirichlet(12, 11)
    n = A.shape[0]
    perm = ct.mmd(A, delta=delta)
    assert sorted(perm.tolist()) == list(range(1, n + 1))
    deg = np.diff(sp.csc_matrix(A).indptr) - 1
    assert deg[perm[0] - 1] == deg.min()
    Ad = A.toarray()
    assert _fill(Ad, perm) < 0.6 * _fill(Ad, np.arange(1, n + 1))
    P, S = ct.symbolic(A, alg=perm)
    check_invariants(ct, sp.csc_matrix(A), P, S)
    # isolated vertices come first, in vertex order
    B = sp.block_diag([sp.identity(3), A, sp.identity(2)]).tocsc()
    pb = ct.mmd(B, delta=delta)
    assert pb[:5].tolist() == [1, 2, 3, n + 4, n + 5]


def test_mmd_random_graphs(ct):
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = int(rng.integers(5, 60))
        M = sp.random(n, n, density=float(rng.uniform(0.02, 0.3)), random_state=int(rng.integers(1 << 30)))
        M = ((M + M.T) != 0).astype(float) + sp.identity(n)
        perm = ct.mmd(M.tocsc())
        assert sorted(perm.tolist()) == list(range(1, n + 1))
        P, S = ct.symbolic(M.tocsc(), alg=perm)
        check_invariants(ct, sp.csc_matrix(M), P, S)
