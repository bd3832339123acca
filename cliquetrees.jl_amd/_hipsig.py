"""ctypes signatures of libmfhip.so (include/mfhip.h)."""
import ctypes

i64 = ctypes.c_int64
i64p = ctypes.POINTER(ctypes.c_int64)
f64p = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p
ci = ctypes.c_int
cd = ctypes.c_double

SIGNATURES = {
    "mf_last_error": ([], ctypes.c_char_p),
    "mf_plan_create": ([i64, i64] + [i64p] * 8 + [i64, ctypes.POINTER(vp)], ci),
    "mf_plan_destroy": ([vp], ci),
    "mf_plan_create_tuned": ([i64, i64] + [i64p] * 8 + [i64p, i64, ctypes.POINTER(vp)], ci),
    "mf_plan_tuning": ([vp, i64p, i64], ci),
    "mf_plan_stats": ([vp, i64p, i64], ci),
    "mf_factor_create": ([vp, ci, ctypes.POINTER(vp)], ci),
    "mf_factor_destroy": ([vp], ci),
    "mf_factor_bytes": ([vp, i64p], ci),
    "mf_factor_set_stream": ([vp, vp], ci),
    "mf_factor_sync": ([vp], ci),
    "mf_factor_set_solve_fp64": ([vp, ci], ci),
    "mf_factor_set_values": ([vp, vp, vp, ci], ci),
    "mf_factor_load": ([vp, vp, vp, vp, ci, ci], ci),
    "mf_factor_set_pattern": ([vp, i64, i64p], ci),
    "mf_factor_refresh": ([vp, vp, ci], ci),
    "mf_factorize": ([vp, ci, vp, ci, cd, cd, i64p], ci),
    "mf_factorize_async": ([vp, ci, vp, ci, cd, cd], ci),
    "mf_factor_info": ([vp, i64p], ci),
    "mf_factor_get": ([vp, vp, vp, vp, ci], ci),
    "mf_solve_natural": ([vp, vp, i64, i64, ci], ci),
    "mf_solve": ([vp, i64p, vp, i64, i64, ci], ci),
    "mf_solve_noalloc": ([vp, i64p, vp, i64, i64], ci),
    "mf_chol_impl": ([vp, f64p, f64p, vp, vp, ci, ci, ci, cd, cd, i64p], ci),
    "mf_selinv": ([vp], ci),
    "mf_selinv_get": ([vp, vp, vp, ci], ci),
    "mf_selinv_impl": ([vp, f64p, f64p, ci, i64p], ci),
    "mf_complete_impl": ([vp, f64p, f64p, ci, i64p], ci),
    "mf_div_impl": ([vp, f64p, f64p, vp, f64p, i64, i64, ci, ci], ci),
    "mf_factor_set_uplo": ([vp, ci], ci),
    "mf_timing": ([vp, ci], ci),
    "mf_timing_read": ([vp, f64p, i64p, f64p, ci], ci),
    "mf_dag_profile": ([vp, ci], ci),
    "mf_factor_set_matrix": ([vp, i64, i64p, i64p, vp, ci], ci),
    "mf_refine": ([vp, i64p, vp, i64, vp, i64, i64, ci, cd, ci, ctypes.POINTER(ci), f64p], ci),
    "mf_refine_segmented": ([vp, i64p, vp, i64, vp, i64, i64, ci, cd, ci, i64, i64p,
                             ctypes.POINTER(ci), f64p], ci),
    "mf_refine_steps": ([vp, i64p, vp, i64, vp, i64, i64, ci, i64, i64p, vp], ci),
    "mf_dag_profile_read": ([vp, f64p, ci], ci),
    "mf_dag_trace_read": ([vp, vp, i64], ci),
    "mf_plan_dag": ([vp, vp, i64], ci),
    "mf_plan_gemm": ([vp, vp, i64], ci),
    "mf_solve_trace_read": ([vp, vp, i64], ci),
    "mf_plan_solve_tasks": ([vp, vp, i64], ci),
    "mf_plan_fronts": ([vp, i64p, i64], ci),
}

# Names of the C-ABI entry points declared in include/mfhip.h.
EXPORTS = tuple(SIGNATURES)


def declare(lib):
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
