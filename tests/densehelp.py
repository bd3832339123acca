"""Dense views of the reference's factor layout, for closed-form checks.

``Dval`` holds, per front j, the nn x nn diagonal block (column-major) at
``Dptr[j]``; ``Lval`` holds the na x nn block below it at ``Lptr[j]``, its
rows being ``sep(j)`` (chordal_factorization.jl:1-21,
cholesky.jl:467-500).  These helpers move between that layout and a dense
n x n lower-triangular matrix in permuted order, so a factor computed by
dense LAPACK on PAP' can be compared with the sparse one entry by entry.
"""
import numpy as np
import scipy.sparse as sp


def _blocks(S):
    nn, na = S.nn(), S.na()
    for j in range(S.nfr):
        r0 = int(S.resptr[j] - 1)
        k, a = int(nn[j]), int(na[j])
        rows = S.septgt[S.sepptr[j] - 1: S.sepptr[j] - 1 + a] - 1
        yield j, r0, k, a, rows


def dense_lower(S, Dval, Lval, unit=False):
    """n x n lower triangle of the factor stored in Dval/Lval (permuted order).
    ``unit`` puts ones on the diagonal (LDLt, whose Dval diagonal is not L's)."""
    n = S.n
    M = np.zeros((n, n))
    for j, r0, k, a, rows in _blocks(S):
        blk = np.asarray(Dval[S.Dptr[j] - 1: S.Dptr[j] - 1 + k * k]).reshape(k, k, order="F")
        M[r0:r0 + k, r0:r0 + k] = np.tril(blk)
        M[np.ix_(rows, np.arange(r0, r0 + k))] = np.asarray(
            Lval[S.Lptr[j] - 1: S.Lptr[j] - 1 + a * k]).reshape(a, k, order="F")
    if unit:
        np.fill_diagonal(M, 1.0)
    return M


def to_layout(S, L):
    """Dval, Lval of a dense lower-triangular n x n factor (permuted order):
    the entries the symbolic pattern stores, the rest dropped."""
    Dval = np.zeros(S.ndz)
    Lval = np.zeros(S.nlz)
    for j, r0, k, a, rows in _blocks(S):
        Dval[S.Dptr[j] - 1: S.Dptr[j] - 1 + k * k] = np.tril(L[r0:r0 + k, r0:r0 + k]).ravel(order="F")
        Lval[S.Lptr[j] - 1: S.Lptr[j] - 1 + a * k] = L[np.ix_(rows, np.arange(r0, r0 + k))].ravel(order="F")
    return Dval, Lval


def outside_pattern_norm(S, L):
    """Frobenius norm of the entries of a dense lower factor that the symbolic
    pattern does not store (zero for an exact factor of PAP')."""
    M = np.tril(np.array(L, dtype=np.float64, copy=True))
    for j, r0, k, a, rows in _blocks(S):
        M[r0:r0 + k, r0:r0 + k] = 0.0
        M[np.ix_(rows, np.arange(r0, r0 + k))] = 0.0
    return float(np.linalg.norm(M))


def permuted_dense(A, P):
    """PAP' as a dense array: (PAP')[i, j] = A[perm[i], perm[j]]."""
    Ad = A.toarray() if sp.issparse(A) else np.asarray(A, dtype=np.float64)
    p = P.perm - 1
    return Ad[np.ix_(p, p)]


def nod_columns(S):
    """Permuted columns (0-based) that form an nn = 1 front."""
    nn = S.nn()
    return np.array([S.resptr[j] - 1 for j in range(S.nfr) if nn[j] == 1], dtype=np.int64)


def rel(a, b):
    """Normwise relative distance, the reference's isapprox measure
    (chordal_triangular.jl:648-660,710-713)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.linalg.norm(a), np.linalg.norm(b), 1e-300)
    return float(np.linalg.norm(a - b) / den)


def first_indefinite_minor(M):
    """1-based k of the first leading principal submatrix M[:k, :k] that is
    not positive definite (its smallest eigenvalue <= 0), 0 if none."""
    for k in range(1, M.shape[0] + 1):
        if np.linalg.eigvalsh(M[:k, :k])[0] <= 0.0:
            return k
    return 0


def leading_pivots(M):
    """Pivots of the unpivoted LDL' of M from ratios of leading principal
    minors, D_k = det(M_k) / det(M_{k-1}) (log-determinants with signs)."""
    out = np.empty(M.shape[0])
    prev_s, prev_l = 1.0, 0.0
    for k in range(1, M.shape[0] + 1):
        s, l = np.linalg.slogdet(M[:k, :k])
        out[k - 1] = s * prev_s * np.exp(l - prev_l)
        prev_s, prev_l = s, l
    return out


# The README examples (src/Multifrontal.jl/README.md:29-44, 89-110, 145-157)
README_A = np.array([[4, 2, 0, 0, 2], [2, 5, 0, 0, 3], [0, 0, 4, 2, 0], [0, 0, 2, 5, 2],
                     [2, 3, 0, 2, 7]], float)
README_L = np.array([[2, 0, 0, 0, 0], [0, 2, 0, 0, 0], [0, 1, 2, 0, 0], [1, 0, 0, 2, 0],
                     [0, 1, 1, 1, 2]], float)
README_AI = np.array([[-4, -2, 0, 0, 2], [-2, -5, 0, 0, 3], [0, 0, 4, 2, 0], [0, 0, 2, 5, 2],
                      [2, 3, 0, 2, 7]], float)
README_LU = np.array([[1, 0, 0, 0, 0], [0, 1, 0, 0, 0], [0, .5, 1, 0, 0], [.5, 0, 0, 1, 0],
                      [0, -.5, -.5, .5, 1]])
README_D = np.array([4.0, -4.0, -4.0, 4.0, 8.0])
README_PERM = [3, 1, 2, 4, 5]
# selinv! of the README matrix, as printed (src/Multifrontal.jl/src/selinv.jl:14-31):
# the lower triangle of (PAP')^-1 on the factor's pattern, perm README_PERM
README_SELINV = np.array([[0.328125, 0, 0, 0, 0],
                          [0.0, 0.328125, 0, 0, 0],
                          [0.0, -0.09375, 0.3125, 0, 0],
                          [-0.15625, 0.0, 0.0, 0.3125, 0],
                          [0.0, -0.0625, -0.125, -0.125, 0.25]])
# elimination orders whose postordered supernodal tree gives README_PERM
README_ORDERS = [[3, 4, 1, 2, 5], [3, 1, 4, 2, 5]]

# DynamicRegularization docstring (dynamic_regularization.jl:12-30): A and
# Matrix(F) as printed (six significant digits)
DYNREG_A = np.array([[1, 1, 2], [1, 1, 3], [2, 3, 1]], float)
DYNREG_B = np.array([[1.0, 1.0, 2.0], [1.0, 1.00001, 3.0], [2.0, 3.0, 1.65144e5]])


def weakly_indefinite(A, bad, seed, coupling=1e-3, value=-10.0):
    """A copy of the SPD matrix A whose rows/columns ``bad`` (0-based) are
    coupled by ``coupling`` times their old entries and carry the diagonal
    ``value``: their pivots are negative, while the growth a replaced pivot
    of size delta causes stays small (coupling^2 / delta)."""
    A = sp.lil_matrix(A, dtype=np.float64)
    rng = np.random.default_rng(seed)
    B = sp.csc_matrix(A)
    scale = np.ones(A.shape[0])
    scale[bad] = coupling
    B = sp.diags(scale) @ B @ sp.diags(scale)
    B = sp.lil_matrix(B)
    for i in bad:
        B[i, i] = value * (1.0 + 0.1 * rng.random())
    return sp.csc_matrix(B)


def to_upper_layout(S, Dval, Lval):
    """Dval/Lval of the :L layout rewritten in the :U layout (each diagonal
    block and each off-diagonal block transposed, chordal_triangular.jl:1015-1055)."""
    D = np.zeros(S.ndz)
    L = np.zeros(S.nlz)
    for j, r0, k, a, rows in _blocks(S):
        p, q = S.Dptr[j] - 1, S.Lptr[j] - 1
        D[p:p + k * k] = np.asarray(Dval[p:p + k * k]).reshape(k, k, order="F").T.ravel(order="F")
        L[q:q + a * k] = np.asarray(Lval[q:q + a * k]).reshape(a, k, order="F").T.ravel(order="F")
    return D, L


def sparse_lower(S, Dval, Lval, unit=False):
    """The factor stored in Dval/Lval as a scipy CSC lower-triangular matrix
    (permuted order); ``unit`` puts ones on the diagonal."""
    rows, cols, vals = [], [], []
    for j, r0, k, a, srow in _blocks(S):
        blk = np.asarray(Dval[S.Dptr[j] - 1: S.Dptr[j] - 1 + k * k]).reshape(k, k, order="F")
        ii, jj = np.tril_indices(k)
        v = blk[ii, jj].copy()
        if unit:
            v[ii == jj] = 1.0
        rows.append(r0 + ii)
        cols.append(r0 + jj)
        vals.append(v)
        if a:
            Lb = np.asarray(Lval[S.Lptr[j] - 1: S.Lptr[j] - 1 + a * k]).reshape(a, k, order="F")
            rr, cc = np.meshgrid(srow, np.arange(r0, r0 + k), indexing="ij")
            rows.append(rr.ravel())
            cols.append(cc.ravel())
            vals.append(Lb.ravel())
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(S.n, S.n))
