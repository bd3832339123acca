"""HIP graph capture of one device-resident step (value refresh,
factorization and 8-RHS solve through the C-ABI) on a torch stream: every
replay, with new values and right-hand sides written into the captured
buffers between replays and eager work in between, equals the eager calls
bit for bit.  The engine clears its counters with a fill kernel, never with
hipMemsetAsync, and allocates its workspaces on the eager warm-up call, so
nothing in the captured sequence allocates or synchronizes."""
import ctypes

import numpy as np
import pytest

from densehelp import rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_matches_eager(ct, oracle, precision):
    import torch
    from cliquetrees_amd import _native
    lib = _native.hip()
    chk = ct.factorization._check
    A = ct.problems.laplacian3d_7pt(16, 16, 16, seed=2)
    P, S = ct.symbolic(A, alg=ct.grid_nd(16, 16, 16))
    print(f"graph replay precision={precision}", flush=True)
    plan = ct.device_plan(S)
    fac = ct.DeviceFactor(plan, precision)
    h = fac.handle
    flat = ct.flatindices(S, A)
    chk(lib.mf_factor_set_pattern(h, len(flat), flat.ctypes.data_as(_native.i64p)))
    stream = torch.cuda.Stream()
    chk(lib.mf_factor_set_stream(h, ctypes.c_void_p(stream.cuda_stream)))
    n, nrhs = S.n, 8
    dev = torch.device("cuda:0")
    nz = torch.tensor(np.ascontiguousarray(A.data), dtype=torch.float64, device=dev)
    B = torch.zeros((nrhs, n), dtype=torch.float64, device=dev)   # column-major n x nrhs
    X = torch.zeros_like(B)
    perm = np.ascontiguousarray(P.perm, dtype=np.int64)

    def step():
        X.copy_(B)
        chk(lib.mf_factor_refresh(h, ctypes.c_void_p(nz.data_ptr()), 1))
        chk(lib.mf_factorize_async(h, 0, None, 0, 0.0, 0.0))
        chk(lib.mf_solve(h, perm.ctypes.data_as(_native.i64p), ctypes.c_void_p(X.data_ptr()), nrhs, n, 1))

    rng = np.random.default_rng(5)
    with torch.cuda.stream(stream):
        step()  # workspaces allocated here
    stream.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        step()
    stream.synchronize()
    for rep in range(4):
        scale = rng.uniform(0.5, 2.0)
        vals = np.ascontiguousarray(A.data) * scale
        Bh = rng.standard_normal((n, nrhs))
        nz.copy_(torch.from_numpy(vals))
        B.copy_(torch.from_numpy(np.ascontiguousarray(Bh.T)))
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            step()
        stream.synchronize()
        Xe = X.clone()
        X.zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            g.replay()
        stream.synchronize()
        info = ctypes.c_int64(-1)
        chk(lib.mf_factor_info(h, ctypes.byref(info)))
        assert info.value == 0
        assert torch.equal(X, Xe), f"replay {rep} differs from the eager step"
        Xh = X.cpu().numpy().T
        tol = 1e-11 if precision == 0 else 1e-4
        assert rel((A * scale) @ Xh, Bh) <= tol
