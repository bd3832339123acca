"""mf_solve_noalloc (the solve without device allocation, its workspace in the
library's fixed device array) against mf_solve on the same factor, through
the C-ABI: two calls in a row, the second with a smaller input, each equal
bit for bit to mf_solve's result and within 1e-11 of the oracle."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(ct, oracle, A, order, nrhs, precision, seed):
    import torch
    from cliquetrees_amd import _native
    lib = _native.hip()
    P, S = ct.symbolic(A, alg=order)
    F = ct.ChordalCholesky(A, P, S, precision=precision)
    ct.cholesky_(F)
    n = A.shape[0]
    h = F._f.handle
    perm = np.ascontiguousarray(P.perm, dtype=np.int64)
    B = np.random.default_rng(seed).standard_normal((n, nrhs))
    Bd = torch.tensor(B.T.copy(), dtype=torch.float64, device="cuda")   # column-major n x nrhs
    X1 = Bd.clone()
    X2 = Bd.clone()
    assert lib.mf_solve(h, perm.ctypes.data_as(_native.i64p), ctypes.c_void_p(X1.data_ptr()),
                        nrhs, n, 1) == 0
    rc = lib.mf_solve_noalloc(h, perm.ctypes.data_as(_native.i64p), ctypes.c_void_p(X2.data_ptr()),
                              nrhs, n)
    assert rc == 0, lib.mf_last_error()
    assert lib.mf_factor_sync(h) == 0
    torch.cuda.synchronize()
    assert torch.equal(X1, X2)
    X = X2.cpu().numpy().T
    if precision == ct.MF_FP64:
        Dv, Lv = oracle.assemble(S, P, A)
        D0, L0, d0, _ = oracle.chol_impl(S, Dv, Lv)
        X0 = oracle.solve(S, P, D0, L0, d0, B)
        assert np.linalg.norm(X - X0) / np.linalg.norm(X0) <= 1e-11
    else:
        assert np.linalg.norm(A @ X - B) / np.linalg.norm(B) <= 1e-3
    return F, h, perm


def test_noalloc_solve_matches_solve(ct, oracle):
    from cliquetrees_amd import _native
    lib = _native.hip()
    pr = ct.problems
    print("noalloc call 1: 7-point 20x18x16, 8 rhs, fp64", flush=True)
    _case(ct, oracle, pr.laplacian3d_7pt(20, 18, 16, seed=3), ct.grid_nd(20, 18, 16), 8, ct.MF_FP64, 1)
    print("noalloc call 2: 5-point 24x24, 3 rhs, fp64", flush=True)
    _case(ct, oracle, pr.laplacian2d_dirichlet(24, 24), ct.grid_nd(24, 24, 1), 3, ct.MF_FP64, 2)
    print("noalloc call 3: 27-point 8x8x8, 2 rhs, fp32", flush=True)
    F, h, perm = _case(ct, oracle, pr.stencil3d_27pt(8, 8, 8, seed=3), ct.grid_nd(8, 8, 8), 2,
                       ct.MF_FP32, 3)
    # a solve larger than the fixed array is refused before anything is enqueued
    huge = (1 << 26) // 4
    rc = lib.mf_solve_noalloc(h, perm.ctypes.data_as(_native.i64p), ctypes.c_void_p(16), huge, F.n)
    assert rc == -1 and b"scratch" in lib.mf_last_error()
