"""The selected-inversion schedule of the engine (mf_selinv_plan.cpp), run on
the host in ticket order through mf_selinv_check: every wait must already be
satisfied when its task runs, every product stays inside its buffers, and the
result equals the oracle's selinv_impl! restatement (itself pinned by the
selinv! docstring and dense inverses)."""
import ctypes

import numpy as np
import pytest

from densehelp import rel


def selinv_check(ct, S, D, L):
    from cliquetrees_amd import _native
    lib = _native.sym()
    arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in
            [S.resptr, S.sepptr, S.septgt, S.reltgt, S.chdptr, S.chdtgt, S.Dptr, S.Lptr]]
    f64p = ctypes.POINTER(ctypes.c_double)
    st = np.zeros(9, dtype=np.int64)
    msg = ctypes.create_string_buffer(256)
    if D is None:  # the schedule and its validation only
        rc = lib.mf_selinv_check(S.n, S.nfr, *[a.ctypes.data_as(_native.i64p) for a in arrs],
                                 None, None, None, None, st.ctypes.data_as(_native.i64p), len(st), msg, 256)
        return rc, msg.value.decode(), None, None, st
    D = np.ascontiguousarray(D, dtype=np.float64)
    L = np.ascontiguousarray(L, dtype=np.float64)
    SD, SL = np.zeros_like(D), np.zeros_like(L)
    rc = lib.mf_selinv_check(S.n, S.nfr, *[a.ctypes.data_as(_native.i64p) for a in arrs],
                             D.ctypes.data_as(f64p), L.ctypes.data_as(f64p), SD.ctypes.data_as(f64p),
                             SL.ctypes.data_as(f64p), st.ctypes.data_as(_native.i64p), len(st), msg, 256)
    return rc, msg.value.decode(), SD, SL, st


def _cases(ct):
    pr = ct.problems
    yield "lap2d_30_nd", pr.laplacian2d_dirichlet(30, 30), ct.grid_nd(30, 30, 1), "maximal"
    yield "lap3d_12_nd", pr.laplacian3d_7pt(12, 12, 12, seed=2), ct.grid_nd(12, 12, 12), "maximal"
    yield "st27_9_nd", pr.stencil3d_27pt(9, 9, 9, seed=3), ct.grid_nd(9, 9, 9), "maximal"
    yield "lap3d_10_mmd", pr.laplacian3d_7pt(10, 10, 10, seed=4), "mmd", "maximal"
    yield "lap3d_8_fund", pr.laplacian3d_7pt(8, 8, 8, seed=5), ct.grid_nd(8, 8, 8), "fundamental"


CASES = ["lap2d_30_nd", "lap3d_12_nd", "st27_9_nd", "lap3d_10_mmd", "lap3d_8_fund"]


@pytest.fixture(scope="module")
def cases(ct):
    return {name: rest for name, *rest in _cases(ct)}


@pytest.mark.parametrize("name", CASES)
def test_selinv_schedule_on_host_matches_oracle(ct, oracle, cases, name):
    A, order, snd = cases[name]
    P, S = ct.symbolic(A, alg=order, snd=snd)
    Dv, Lv = oracle.assemble(S, P, A)
    D, L, d, info = oracle.chol_impl(S, Dv, Lv)
    assert info == 0
    rc, msg, SD, SL, st = selinv_check(ct, S, D, L)
    assert rc == 0, msg
    OD, OL = oracle.selinv(S, D, L)
    assert rel(SD, OD) <= 1e-12 and rel(SL, OL) <= 1e-12


def test_selinv_schedule_big_fronts(ct, oracle):
    """A 3-D grid whose top separators give big fronts with several 64-column
    tiles (the block recursion of X = L11^-1 over more than one level)."""
    A = ct.problems.laplacian3d_7pt(16, 16, 16, seed=7)
    P, S = ct.symbolic(A, alg=ct.grid_nd(16, 16, 16))
    assert int(np.diff(S.resptr).max()) > 128
    Dv, Lv = oracle.assemble(S, P, A)
    D, L, d, info = oracle.chol_impl(S, Dv, Lv)
    rc, msg, SD, SL, st = selinv_check(ct, S, D, L)
    assert rc == 0, msg
    OD, OL = oracle.selinv(S, D, L)
    assert rel(SD, OD) <= 1e-12 and rel(SL, OL) <= 1e-12
    assert st[1] > 0 and st[3] > 0


def test_selinv_schedule_config_sizes(ct):
    """The schedule of the 64^3 7-point Laplacian builds and validates (bounds
    of every product; no execution); its work is about twice the factor's."""
    A = ct.problems.laplacian3d_7pt(64, 64, 64, seed=2)
    P, S = ct.symbolic(A, alg=ct.grid_nd(64, 64, 64))
    rc, msg, _, _, st = selinv_check(ct, S, None, None)
    assert rc == 0, msg
    fl = S.factor_flops()
    assert 1.5 * fl < st[4] < 3.0 * fl
    # X, Lt and M of FP64 fit well inside one device
    assert 8 * (2 * st[5] + st[6] + st[7]) < 32e9
