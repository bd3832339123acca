"""The completion's backward-solve schedule (build_cpl_solve,
mf_selinv_plan.cpp), run on the host in ticket order through
mf_cpl_solve_check: every wait is met when its task runs, no task reads a
value that no earlier task wrote, and the rows the completion reads equal
backward substitution with the factor (L^-T e, a textbook computation)."""
import ctypes

import numpy as np
import pytest


def cpl_check(na, nn, seed=1):
    from cliquetrees_amd import _native
    lib = _native.sym()
    na = np.ascontiguousarray(na, dtype=np.int64)
    nn = np.ascontiguousarray(nn, dtype=np.int64)
    worst = ctypes.c_double(-1.0)
    st = np.zeros(3, dtype=np.int64)
    msg = ctypes.create_string_buffer(256)
    rc = lib.mf_cpl_solve_check(len(na), na.ctypes.data_as(_native.i64p), nn.ctypes.data_as(_native.i64p),
                                seed, ctypes.byref(worst), st.ctypes.data_as(_native.i64p), len(st), msg, 256)
    return rc, msg.value.decode(), worst.value, st


@pytest.mark.parametrize("seed", [1, 2])
def test_cpl_solve_schedule_mixed_fronts(ct, seed):
    """Small and big roots, separators on and off tile boundaries, residuals
    of one tile, several tiles and a partial last tile."""
    na = [100, 0, 70, 10, 200, 63, 129, 1, 64, 0, 5]
    nn = [5, 130, 70, 3, 64, 2, 129, 64, 65, 40, 200]
    rc, msg, worst, st = cpl_check(na, nn, seed)
    assert rc == 0, msg
    assert worst <= 1e-12
    assert st[0] > 0 and st[2] > 0


def test_cpl_solve_schedule_long_chains(ct):
    """Chains longer than the look-ahead (the reduction split between the
    chain and a task of its own): a wide separator, a wide residual, both."""
    rc, msg, worst, st = cpl_check([700, 0, 530, 1000], [70, 600, 300, 1])
    assert rc == 0, msg
    assert worst <= 1e-12


def test_cpl_solve_schedule_of_a_grid(ct):
    """The fronts of a 12^3 nested-dissection Laplacian (the sizes the
    completion of that factor builds)."""
    A = ct.problems.laplacian3d_7pt(12, 12, 12, seed=2)
    P, S = ct.symbolic(A, alg=ct.grid_nd(12, 12, 12))
    rc, msg, worst, st = cpl_check(np.diff(S.sepptr), np.diff(S.resptr))
    assert rc == 0, msg
    assert worst <= 1e-12


def test_cpl_solve_schedule_only_small(ct):
    """Only small roots: no tasks."""
    rc, msg, worst, st = cpl_check([3, 0, 60], [2, 64, 4])
    assert rc == 0, msg
    assert st[0] == 0
