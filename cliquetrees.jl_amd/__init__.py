"""MI355X-native numeric engine for the multifrontal path of CliqueTrees.jl.

The directory name is not a Python identifier; load it with
``importlib.util.spec_from_file_location`` (see ``load_package`` in
``__graft_entry__.py``) under the module name ``cliquetrees_amd``.
"""
from .symbolic import ChordalSymbolic, Permutation, symbolic, flatindices, grid_nd, mmd  # noqa: F401
from .factorization import (  # noqa: F401
    ChordalCholesky, ChordalLDLt, ChordalFactorization, DynamicRegularization, GMW81,
    PosDefException, SingularException, MFError, cholesky_, ldlt_, ldiv_, solve, rdiv_, rsolve, issuccess, refine,
    chol_impl, div_impl, selinv_, selinv_impl, complete_, complete_impl, device_plan, DevicePlan, DeviceFactor, checkinfo, TUNING_NAMES,
    save_factor, load_factor, FACTOR_FORMAT,
    MF_FP64, MF_FP32, MF_CHOLESKY, MF_LDLT, MF_UPLO_L, MF_UPLO_U, STAT_NAMES, KC_NAMES,
)
from . import problems  # noqa: F401
from .batched import BatchedFactorization, BatchFailure, batched_cholesky, batched_ldlt  # noqa: F401
from . import serving  # noqa: F401
from . import autotune  # noqa: F401
