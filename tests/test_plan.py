"""The numeric engine's planner, run on the host through mf_plan_check:
internal fronts, schedule, memory layout and the bounds of every task."""
import ctypes

import numpy as np
import pytest


def plan_check(ct, S):
    from cliquetrees_amd import _native
    lib = _native.sym()
    arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in
            [S.resptr, S.sepptr, S.septgt, S.reltgt, S.chdptr, S.chdtgt, S.Dptr, S.Lptr]]
    st = np.zeros(19, dtype=np.int64)
    msg = ctypes.create_string_buffer(256)
    rc = lib.mf_plan_check(S.n, S.nfr, *[a.ctypes.data_as(_native.i64p) for a in arrs],
                           st.ctypes.data_as(_native.i64p), len(st), msg, 256)
    return rc, msg.value.decode(), st


def test_plan_small_problems(ct):
    P = ct.problems
    probs = [(P.laplacian2d_dirichlet(30, 30), ct.grid_nd(30, 30, 1)),
             (P.laplacian2d_dirichlet(20, 20), None),
             (P.laplacian3d_7pt(16, 16, 16, seed=2), ct.grid_nd(16, 16, 16)),
             (P.stencil3d_27pt(10, 10, 10, seed=3), ct.grid_nd(10, 10, 10)),
             (np.eye(7), None)]
    for A, order in probs:
        for snd in ["maximal", "nodal", "fundamental"]:
            Pm, S = ct.symbolic(A, alg=order, snd=snd)
            rc, msg, st = plan_check(ct, S)
            assert rc == 0, msg
            # the engine's flops include the explicit zeros of merged chains only
            assert st[11] >= st[10] * (1 - 1e-9)


def test_plan_config_sizes(ct):
    """Full-size 3-D configurations: 64^3 7-point and 48^3 27-point."""
    for A, dims in [(ct.problems.laplacian3d_7pt(64, 64, 64, seed=2), (64, 64, 64)),
                    (ct.problems.stencil3d_27pt(48, 48, 48, seed=3), (48, 48, 48))]:
        Pm, S = ct.symbolic(A, alg=ct.grid_nd(*dims))
        rc, msg, st = plan_check(ct, S)
        assert rc == 0, msg
        # merged chains add at most a few percent of explicit-zero work
        assert st[11] <= 1.1 * st[10]
