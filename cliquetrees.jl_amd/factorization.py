"""Host mirror of the reference's factorization API for the multifrontal path.

Names follow src/Multifrontal.jl (Python has no ``!``: the in-place verbs
carry a trailing underscore):

  ``ChordalCholesky(A[, P, S])`` / ``ChordalLDLt(A[, P, S])``
        chordal_factorization.jl:45-121,197-244: symbolic analysis, then the
        values of A are assembled (``copyto!``, chordal_triangular.jl:852-893)
        on the device from ``flatindices`` (chordal_symbolic.jl:595-752).
  ``cholesky_(F; check, reg)``   cholesky!  cholesky.jl:80-101
  ``ldlt_(F; signs, reg, check)`` ldlt!     cholesky.jl:181-203
  ``ldiv_(F, B)`` / ``solve(F, B)`` (= ``F \\ B``)  ldiv! divide.jl:201-206
  ``rdiv_(B, F)`` / ``rsolve(B, F)`` (= ``B / F``)   rdiv! divide.jl:71-74,277-285
  ``issuccess(F)``               abstract_factorization.jl:100-102
  ``chol_impl(S, Dval, Lval, ...)`` / ``div_impl(...)``: the flat-array seam
        (cholesky.jl:361-392, divide.jl:417-450) on host arrays.

Every numeric call goes to libmfhip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import scipy.sparse as sp

from . import _native
from .symbolic import ChordalSymbolic, Permutation, flatindices, symbolic

MF_FP64, MF_FP32 = 0, 1
MF_CHOLESKY, MF_LDLT = 0, 1
MF_REG_NONE, MF_REG_DYNAMIC, MF_REG_GMW81 = 0, 1, 2
MF_UPLO_L, MF_UPLO_U = 0, 1
_UPLO = {"L": MF_UPLO_L, "U": MF_UPLO_U}

STAT_NAMES = ["n", "nfr", "internal_fronts", "small_region", "big_fronts", "big_levels",
              "small_levels", "max_nj", "panel_elems", "u_elems", "w_elems", "gemm_tasks",
              "ea_tasks", "diag_tasks", "ref_flops", "engine_flops", "gemm_flops", "ndz", "nlz",
              "dag_tasks", "ref_flops_dag", "ref_flops_small", "mv_elems1"]
KC_NAMES = ["small", "ea", "diag", "gemm", "copyout", "solve", "assemble", "dag", "selinv"]


class PosDefException(Exception):
    """Mirror of LinearAlgebra.PosDefException (utils.jl:778-779)."""

    def __init__(self, info):
        super().__init__(f"matrix is not positive definite; factorization failed at column {info}")
        self.info = info


class SingularException(Exception):
    """Mirror of LinearAlgebra.SingularException (utils.jl:780-781)."""

    def __init__(self, info):
        super().__init__(f"matrix is singular; zero pivot at column {info}")
        self.info = info


class MFError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        msg = _native.hip().mf_last_error()
        raise MFError(f"libmfhip error {rc}: {msg.decode() if msg else ''}")


def checkinfo(info, diag):
    """checkinfo (utils.jl:775-785)."""
    if info < 0:
        raise ValueError(info)
    if info > 0:
        raise (PosDefException if diag == MF_CHOLESKY else SingularException)(info)


class DynamicRegularization:
    """DynamicRegularization (dynamic_regularization.jl:38-90): pivots with
    sign*Djj < epsilon are replaced by delta*sign; negative parameters take the
    defaults cbrt(eps(T)) and sqrt(eps(T))."""

    def __init__(self, delta=-1.0, epsilon=-1.0):
        self.delta = delta
        self.epsilon = epsilon

    def initialize(self, eps):
        d = self.delta if self.delta >= 0 else eps ** (1.0 / 3.0)
        e = self.epsilon if self.epsilon >= 0 else math.sqrt(eps)
        return d, e


class GMW81:
    """GMW81 (regularization/gmw81.jl:1-155), the Gill-Murray-Wright 1981
    modified Cholesky: each pivot becomes s * max(|Djj|, Aimax^2 / beta, delta)
    with Aimax the largest magnitude below it in its column.  Negative
    parameters take the defaults: beta = gmw81_beta of the matrix (computed
    on the device), delta = cbrt(eps(T))."""

    def __init__(self, beta=-1.0, delta=-1.0):
        self.beta = beta
        self.delta = delta


def _reg_args(reg):
    """(kind, delta, eps) of mf_factorize for a regularization object."""
    if reg is None:
        return MF_REG_NONE, 0.0, 0.0
    if isinstance(reg, GMW81):
        return MF_REG_GMW81, float(reg.delta), float(reg.beta)
    delta, eps = reg.initialize(np.finfo(np.float64).eps)
    return MF_REG_DYNAMIC, delta, eps


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


TUNING_NAMES = ["upd_group", "ea_rows", "wide_k", "sim_workers", "zero_ppm", "prio", "fuse",
                "wg_per_cu"]


class DevicePlan:
    """A ChordalSymbolic uploaded once (mf_plan_create; with ``tuning``, a
    sequence of the planner parameters TUNING_NAMES, mf_plan_create_tuned)."""

    def __init__(self, S: ChordalSymbolic, tuning=None):
        lib = _native.hip()
        h = ctypes.c_void_p()
        arrs = [S.resptr, S.sepptr, S.septgt, S.reltgt, S.chdptr, S.chdtgt, S.Dptr, S.Lptr]
        self._keep = [np.ascontiguousarray(a, dtype=np.int64) for a in arrs]
        ptrs = [a.ctypes.data_as(_native.i64p) for a in self._keep]
        if tuning is None:
            _check(lib.mf_plan_create(S.n, S.nfr, *ptrs, 0, ctypes.byref(h)))
        else:
            t = np.ascontiguousarray(tuning, dtype=np.int64)
            _check(lib.mf_plan_create_tuned(S.n, S.nfr, *ptrs, t.ctypes.data_as(_native.i64p), len(t),
                                            ctypes.byref(h)))
        self.handle = h
        self.S = S

    def tuning(self):
        """The planner parameters in use and the scheduler's launch estimate (us)."""
        out = np.zeros(len(TUNING_NAMES) + 1, dtype=np.int64)
        _check(_native.hip().mf_plan_tuning(self.handle, out.ctypes.data_as(_native.i64p), len(out)))
        d = dict(zip(TUNING_NAMES, out[:-1].tolist()))
        d["sim_us"] = int(out[-1])
        return d

    def stats(self):
        out = np.zeros(len(STAT_NAMES), dtype=np.int64)
        _check(_native.hip().mf_plan_stats(self.handle, out.ctypes.data_as(_native.i64p), len(out)))
        return dict(zip(STAT_NAMES, out.tolist()))

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _native.hip().mf_plan_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


def device_plan(S: ChordalSymbolic, tuning=None) -> DevicePlan:
    """The device plan of S (one per planner tuning, kept with S)."""
    key = None if tuning is None else tuple(int(x) for x in tuning)
    plans = getattr(S, "_devplans", None)
    if plans is None:
        plans = {}
        S._devplans = plans
    dp = plans.get(key)
    if dp is None:
        dp = DevicePlan(S, key)
        plans[key] = dp
    return dp


class DeviceFactor:
    """Device storage of one numeric factorization (mf_factor_create)."""

    def __init__(self, plan: DevicePlan, precision=MF_FP64):
        h = ctypes.c_void_p()
        _check(_native.hip().mf_factor_create(plan.handle, precision, ctypes.byref(h)))
        self.handle = h
        self.plan = plan
        self.precision = precision

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _native.hip().mf_factor_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


class ChordalFactorization:
    """ChordalFactorization{DIAG, UPLO} (chordal_factorization.jl:1-21) whose
    values live on the GPU.  ``uplo`` ("L" or "U") is the stored triangle of
    ``Dval``/``Lval``: "U" holds the transposed blocks (U = L',
    chordal_triangular.jl:895-926, 1015-1055)."""

    DIAG = MF_CHOLESKY

    def __init__(self, A, P: Permutation | None = None, S: ChordalSymbolic | None = None, *,
                 alg=None, snd="maximal", precision=MF_FP64, uplo="L", tuning=None,
                 solve_fp64=False):
        A = sp.csc_matrix(A, dtype=np.float64)
        A.sort_indices()
        if P is None or S is None:
            P, S = symbolic(A, alg=alg, snd=snd)
        self.P, self.S = P, S
        self.perm, self.invp = P.perm, P.invp
        self.n = S.n
        self.info = 0
        if isinstance(tuning, str):
            if tuning != "auto":
                raise ValueError("tuning must be a sequence of planner parameters or 'auto'")
            from .autotune import lookup
            tuning = lookup(S, precision)
        self._plan = device_plan(S, tuning)
        self._f = DeviceFactor(self._plan, precision)
        self._set_uplo(uplo)
        if solve_fp64:
            # FP32 fronts, solves in FP64 arithmetic on a widened copy of the
            # factor (mf_factor_set_solve_fp64)
            _check(_native.hip().mf_factor_set_solve_fp64(self._f.handle, 1))
        self._pattern = None
        self._matrix_key = None
        self.copyto_(A)

    def _set_uplo(self, uplo):
        if uplo not in _UPLO:
            raise ValueError("uplo must be 'L' or 'U'")
        self.uplo = uplo
        _check(_native.hip().mf_factor_set_uplo(self._f.handle, _UPLO[uplo]))

    # copyto!(F, A) with a pattern reuse fast path (flatindices + setflatindex!)
    def copyto_(self, A):
        A = sp.csc_matrix(A, dtype=np.float64)
        A.sort_indices()
        key = (A.shape, A.indptr.tobytes(), A.indices.tobytes())
        lib = _native.hip()
        if self._pattern != key:
            flat = flatindices(self.S, A, self.uplo)
            _check(lib.mf_factor_set_pattern(self._f.handle, len(flat),
                                             flat.ctypes.data_as(_native.i64p)))
            self._pattern = key
        vals = np.ascontiguousarray(A.data, dtype=np.float64)
        _check(lib.mf_factor_refresh(self._f.handle, _ptr(vals), 0))
        self._A = A
        self.info = 0
        self._selinv = False
        return self

    def set_values(self, Dval, Lval):
        """Load Dval/Lval in the reference layout (e.g. an assembled copyto! result)."""
        Dval = np.ascontiguousarray(Dval, dtype=np.float64)
        Lval = np.ascontiguousarray(Lval, dtype=np.float64)
        _check(_native.hip().mf_factor_set_values(self._f.handle, _ptr(Dval), _ptr(Lval), 0))
        self._selinv = False
        return self

    @property
    def Dval(self):
        out = np.empty(self.S.ndz)
        if getattr(self, "_selinv", False):
            _check(_native.hip().mf_selinv_get(self._f.handle, _ptr(out), None, 0))
        else:
            _check(_native.hip().mf_factor_get(self._f.handle, _ptr(out), None, None, 0))
        return out

    @property
    def Lval(self):
        out = np.empty(self.S.nlz)
        if getattr(self, "_selinv", False):
            _check(_native.hip().mf_selinv_get(self._f.handle, None, _ptr(out), 0))
        else:
            _check(_native.hip().mf_factor_get(self._f.handle, None, _ptr(out), None, 0))
        return out

    @property
    def d(self):
        out = np.empty(self.n)
        _check(_native.hip().mf_factor_get(self._f.handle, None, None, _ptr(out), 0))
        return out

    def stats(self):
        return self._plan.stats()

    @classmethod
    def from_factor(cls, P: Permutation, S: ChordalSymbolic, Dval, Lval, d=None, info=0, uplo="L"):
        """A factorization given by its values, as the reference constructs a
        ChordalFactorization from stored Dval/Lval/d (chordal_factorization.jl:1-21):
        ready for ldiv_/solve without refactoring (mf_factor_load)."""
        F = cls.__new__(cls)
        F.P, F.S = P, S
        F.perm, F.invp = P.perm, P.invp
        F.n = S.n
        F.info = int(info)
        F._plan = device_plan(S)
        F._f = DeviceFactor(F._plan, MF_FP64)
        F._set_uplo(uplo)
        F._pattern = None
        F._matrix_key = None
        F._A = None
        Dval = np.ascontiguousarray(Dval, dtype=np.float64)
        Lval = np.ascontiguousarray(Lval, dtype=np.float64)
        if Dval.shape != (S.ndz,) or Lval.shape != (S.nlz,):
            raise ValueError("Dval/Lval do not match the symbolic factorization")
        dp = None
        if d is not None:
            d = np.ascontiguousarray(d, dtype=np.float64)
            if d.shape != (S.n,):
                raise ValueError("d has the wrong length")
            dp = _ptr(d)
        _check(_native.hip().mf_factor_load(F._f.handle, _ptr(Dval), _ptr(Lval), dp, cls.DIAG, 0))
        return F


class ChordalCholesky(ChordalFactorization):
    DIAG = MF_CHOLESKY


class ChordalLDLt(ChordalFactorization):
    DIAG = MF_LDLT


def _factorize(F, diag, signs, reg, check):
    """factorize! levels 1-3 (cholesky.jl:233-315)."""
    if signs is not None:
        s = np.asarray(signs, dtype=np.float64)
        if s.shape != (F.n,):
            raise ValueError("signs has the wrong length")
        s = np.ascontiguousarray(s[F.perm - 1])  # permuteto (utils.jl:1-9)
        bad = ~((s == 0) | (s == 1) | (s == -1)) if reg is None else ~((s == 1) | (s == -1))
        if bad.any():  # checksigns (no_regularization.jl:7-25)
            raise ValueError("invalid signs")
        sp_ = _ptr(s)
    else:
        sp_ = None
    kind, delta, eps = _reg_args(reg)
    if kind == MF_REG_GMW81 and diag == MF_LDLT and sp_ is None:
        raise ValueError("GMW81 needs signs (checksigns)")
    info = ctypes.c_int64(0)
    F._selinv = False
    _check(_native.hip().mf_factorize(F._f.handle, diag, sp_, kind, delta, eps, ctypes.byref(info)))
    info = int(info.value)
    if check:
        checkinfo(info, diag)
    F.info = 0 if info == 0 else int(F.perm[info - 1])
    return F


def cholesky_(F: ChordalFactorization, check=True, reg=None):
    """cholesky!(F; check, reg) with signs = Ones."""
    if F.DIAG != MF_CHOLESKY:
        raise TypeError("cholesky_ needs a ChordalCholesky")
    return _factorize(F, MF_CHOLESKY, None, reg, check)


def ldlt_(F: ChordalFactorization, signs=None, reg=None, check=True):
    """ldlt!(F; signs, reg, check); default signs = Zeros (no sign check)."""
    if F.DIAG != MF_LDLT:
        raise TypeError("ldlt_ needs a ChordalLDLt")
    return _factorize(F, MF_LDLT, signs, reg, check)


def selinv_(F: ChordalFactorization):
    """selinv!(F::ChordalCholesky) (src/Multifrontal.jl/src/selinv.jl:38-41):
    the selected inverse of the factored matrix, the entries of (P A P')^-1
    on the factor's pattern, computed on the device (mf_selinv).  Afterwards
    F.Dval / F.Lval hold it in F's layout, as the reference's F does; the
    factor itself stays on the device, and the next copyto_/factorization
    replaces the selected inverse."""
    if F.DIAG != MF_CHOLESKY:
        raise TypeError("selinv! is defined for ChordalCholesky factorizations")
    _check(_native.hip().mf_selinv(F._f.handle))
    F._selinv = True
    F.info = 0
    return F


def selinv_impl(S: ChordalSymbolic, Dval, Lval, uplo="L"):
    """selinv_impl! on host arrays through mf_selinv_impl (selinv.jl:72-94):
    Dval/Lval hold a Cholesky factor in the reference layout; returns copies
    holding the selected inverse."""
    Dval = np.array(Dval, dtype=np.float64, copy=True)
    Lval = np.array(Lval, dtype=np.float64, copy=True)
    info = ctypes.c_int64(0)
    plan = device_plan(S)
    _check(_native.hip().mf_selinv_impl(plan.handle, Dval.ctypes.data_as(_native.f64p),
                                        Lval.ctypes.data_as(_native.f64p), _UPLO[uplo],
                                        ctypes.byref(info)))
    return Dval, Lval


def complete_impl(S: ChordalSymbolic, Dval, Lval, uplo="L"):
    """complete_impl! on host arrays through mf_complete_impl
    (complete.jl:46-72): Dval/Lval hold a matrix Sigma on the factor's
    pattern in the reference layout (uplo); returns (Dval, Lval, info) with
    the Cholesky factor of the inverse of Sigma's maximum-determinant
    completion.  info > 0: the local column of the first failing potrf! in
    the reference's order of fronts; the arrays are then returned as given."""
    Dval = np.array(Dval, dtype=np.float64, copy=True)
    Lval = np.array(Lval, dtype=np.float64, copy=True)
    if Dval.shape != (S.ndz,) or Lval.shape != (S.nlz,):
        raise ValueError("Dval/Lval do not match the symbolic factorization")
    info = ctypes.c_int64(0)
    plan = device_plan(S)
    _check(_native.hip().mf_complete_impl(plan.handle, Dval.ctypes.data_as(_native.f64p),
                                          Lval.ctypes.data_as(_native.f64p), _UPLO[uplo],
                                          ctypes.byref(info)))
    return Dval, Lval, int(info.value)


def complete_(F: ChordalFactorization, check=True):
    """complete!(F; check) (complete.jl:1-23): F holds a matrix on its
    pattern (after selinv_, its selected inverse); afterwards F is the
    Cholesky factor of the inverse of that matrix's maximum-determinant
    completion, ready for ldiv_/solve (mf_factor_load), and F.info is set;
    check raises PosDefException on a failed potrf! (checkinfo)."""
    if F.DIAG != MF_CHOLESKY:
        raise TypeError("complete! is defined for ChordalCholesky factorizations")
    D, L, info = complete_impl(F.S, F.Dval, F.Lval, F.uplo)
    F.info = info
    if check:
        checkinfo(info, MF_CHOLESKY)
    if info == 0:
        _check(_native.hip().mf_factor_load(F._f.handle, _ptr(D), _ptr(L), None, MF_CHOLESKY, 0))
        F._selinv = False
    return F


def issuccess(F):
    return F.info == 0


def ldiv_(F: ChordalFactorization, B):
    """ldiv!(F, B): B <- A \\ B in place (B float64, n or n x nrhs)."""
    if not isinstance(B, np.ndarray) or B.dtype != np.float64:
        raise TypeError("B must be a float64 numpy array")
    vec = B.ndim == 1
    M = B.reshape(-1, 1) if vec else B
    if M.shape[0] != F.n:
        raise ValueError("dimension mismatch")
    W = np.asfortranarray(M)
    perm = np.ascontiguousarray(F.perm, dtype=np.int64)
    _check(_native.hip().mf_solve(F._f.handle, perm.ctypes.data_as(_native.i64p), _ptr(W),
                                  W.shape[1], F.n, 0))
    if W is not M:
        M[...] = W
    return B


def solve(F: ChordalFactorization, B):
    """F \\ B."""
    X = np.array(B, dtype=np.float64, copy=True)
    return ldiv_(F, X)


def rdiv_(B, F: ChordalFactorization):
    """rdiv!(B, F): B <- B / F in place (B float64, n or m x n).  A is
    symmetric, so B A^-1 = (A^-1 B')' and the right division is the left
    solve of the transpose (rdiv!, divide.jl:277-285)."""
    if not isinstance(B, np.ndarray) or B.dtype != np.float64:
        raise TypeError("B must be a float64 numpy array")
    if B.shape[-1] != F.n:
        raise ValueError("dimension mismatch")
    if B.ndim == 1:
        return ldiv_(F, B)
    T = np.asfortranarray(B.T)
    ldiv_(F, T)
    B[...] = T.T
    return B


def rsolve(B, F: ChordalFactorization):
    """B / F."""
    X = np.array(B, dtype=np.float64, copy=True)
    return rdiv_(X, F)


def refine(F: ChordalFactorization, B, tol=1e-10, maxit=20, segments=None):
    """X = A \\ B by iterative refinement around F (typically FP32 fronts): the
    residual B - A X and the update are FP64 on the GPU.  Returns X (float64,
    shape of B); F.ir_iters / F.ir_relres hold the corrections done and the
    final worst relative residual.  segments (0-based row offsets, first 0,
    last n) makes the stopping test hold per stacked system.  Raises MFError
    if tol is not reached."""
    lib = _native.hip()
    A = F._A
    key = F._pattern
    if F._matrix_key != key:
        colptr = np.ascontiguousarray(A.indptr, dtype=np.int64) + 1
        rowval = np.ascontiguousarray(A.indices, dtype=np.int64) + 1
        vals = np.ascontiguousarray(A.data, dtype=np.float64)
        _check(lib.mf_factor_set_matrix(F._f.handle, F.n, colptr.ctypes.data_as(_native.i64p),
                                        rowval.ctypes.data_as(_native.i64p), _ptr(vals), 0))
        F._matrix_key = key
    Bm = np.asarray(B, dtype=np.float64)
    vec = Bm.ndim == 1
    Bf = np.asfortranarray(Bm.reshape(-1, 1) if vec else Bm)
    if Bf.shape[0] != F.n:
        raise ValueError("dimension mismatch")
    X = np.empty_like(Bf, order="F")
    perm = np.ascontiguousarray(F.perm, dtype=np.int64)
    iters, rel = ctypes.c_int(0), ctypes.c_double(0.0)
    seg = np.ascontiguousarray([0, F.n] if segments is None else segments, dtype=np.int64)
    _check(lib.mf_refine_segmented(F._f.handle, perm.ctypes.data_as(_native.i64p), _ptr(Bf), F.n,
                                   _ptr(X), F.n, Bf.shape[1], 0, float(tol), int(maxit),
                                   len(seg) - 1, seg.ctypes.data_as(_native.i64p),
                                   ctypes.byref(iters), ctypes.byref(rel)))
    F.ir_iters, F.ir_relres = int(iters.value), float(rel.value)
    if not (F.ir_relres <= tol):
        raise MFError(f"refinement stopped at relative residual {F.ir_relres:.3e} > {tol:.1e}")
    return X.reshape(-1) if vec else X


def chol_impl(S: ChordalSymbolic, Dval, Lval, diag=MF_CHOLESKY, signs=None, reg=None, uplo="L"):
    """chol_impl! on host arrays through mf_chol_impl: returns (Dval, Lval, d, info).
    For Cholesky d stays ones (the reference's d is Ones) and no d is passed."""
    Dval = np.array(Dval, dtype=np.float64, copy=True)
    Lval = np.array(Lval, dtype=np.float64, copy=True)
    d = np.ones(S.n) if diag == MF_CHOLESKY else np.zeros(S.n)
    sp_ = None
    if signs is not None:
        signs = np.ascontiguousarray(signs, dtype=np.float64)
        sp_ = _ptr(signs)
    kind, delta, eps = _reg_args(reg)
    info = ctypes.c_int64(0)
    plan = device_plan(S)
    dp = None if diag == MF_CHOLESKY else _ptr(d)
    _check(_native.hip().mf_chol_impl(plan.handle, Dval.ctypes.data_as(_native.f64p),
                                      Lval.ctypes.data_as(_native.f64p), dp, sp_, diag, _UPLO[uplo],
                                      kind, delta, eps, ctypes.byref(info)))
    return Dval, Lval, d, int(info.value)


def div_impl(S: ChordalSymbolic, Dval, Lval, d, C, diag=MF_CHOLESKY, uplo="L"):
    """Natural solve C <- L^-T [D^-1] L^-1 C on host arrays (mf_div_impl)."""
    C = np.array(C, dtype=np.float64, order="F", copy=True)
    vec = C.ndim == 1
    if vec:
        C = C.reshape(-1, 1, order="F")
    Dval = np.ascontiguousarray(Dval, dtype=np.float64)
    Lval = np.ascontiguousarray(Lval, dtype=np.float64)
    dp = None
    if d is not None:
        d = np.ascontiguousarray(d, dtype=np.float64)
        dp = _ptr(d)
    plan = device_plan(S)
    _check(_native.hip().mf_div_impl(plan.handle, Dval.ctypes.data_as(_native.f64p),
                                     Lval.ctypes.data_as(_native.f64p), dp,
                                     C.ctypes.data_as(_native.f64p), C.shape[1], S.n, diag,
                                     _UPLO[uplo]))
    return C[:, 0] if vec else C


# ---- factor files ---------------------------------------------------------------
# One .npz of flat arrays (no pickled objects): the ChordalSymbolic arrays
# (chordal_symbolic.jl:15-28), the permutation, and the factor values in the
# reference layout (Dval, Lval, d; chordal_factorization.jl:1-21).  Loading
# uploads the plan and the values and inverts the diagonal blocks; the
# factor can be solved with at once.
FACTOR_FORMAT = "cliquetrees-amd-factor-1"
_SYM_ARRAYS = ["resptr", "sepptr", "septgt", "reltgt", "chdptr", "chdtgt", "pnt", "idx", "Dptr", "Lptr"]
_SYM_SCALARS = ["n", "nfr", "nMptr", "nMval", "nNval", "nFval"]


def save_factor(F: ChordalFactorization, path):
    """Write the factor F (after cholesky_/ldlt_ or from_factor) to path (.npz)."""
    if F.info != 0:
        raise ValueError("the factorization failed; nothing to save")
    S = F.S
    out = {"format": np.array(FACTOR_FORMAT), "diag": np.int64(F.DIAG), "info": np.int64(F.info),
           "uplo": np.array(F.uplo),
           "perm": np.asarray(F.perm, dtype=np.int64), "invp": np.asarray(F.invp, dtype=np.int64),
           "Dval": F.Dval, "Lval": F.Lval, "d": F.d}
    for k in _SYM_ARRAYS:
        out[k] = np.asarray(getattr(S, k), dtype=np.int64)
    for k in _SYM_SCALARS:
        out[k] = np.int64(getattr(S, k))
    np.savez(path, **out)


def load_factor(path) -> ChordalFactorization:
    """Read a factor written by save_factor: a ChordalCholesky or ChordalLDLt on
    the device, ready for ldiv_/solve."""
    with np.load(path, allow_pickle=False) as z:
        if str(z["format"]) != FACTOR_FORMAT:
            raise ValueError(f"not a factor file of format {FACTOR_FORMAT}")
        S = ChordalSymbolic(**{k: np.array(z[k], dtype=np.int64) for k in _SYM_ARRAYS},
                            **{k: int(z[k]) for k in _SYM_SCALARS})
        perm = np.array(z["perm"], dtype=np.int64)
        invp = np.array(z["invp"], dtype=np.int64)
        P = Permutation.from_perm(perm)
        if not np.array_equal(P.invp, invp):
            raise ValueError("inconsistent permutation in the factor file")
        diag = int(z["diag"])
        cls = ChordalLDLt if diag == MF_LDLT else ChordalCholesky
        uplo = str(z["uplo"]) if "uplo" in z.files else "L"
        return cls.from_factor(P, S, z["Dval"], z["Lval"], z["d"], info=int(z["info"]), uplo=uplo)

