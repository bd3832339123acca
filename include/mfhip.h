/*
 * mfhip.h — C-ABI of libmfhip.so, the MI355X (gfx950) numeric engine for the
 * multifrontal Cholesky / LDLt path of CliqueTrees.jl.
 *
 * The boundary is the reference's own flat-array seam: `chol_impl!(Mptr,
 * Mval, Dptr, Dval, Lptr, Lval, d, Fval, res, rel, chd, uplo, S, R, diag)`
 * (src/Multifrontal.jl/src/cholesky.jl:361-377) and `div_impl!(B, Mptr, Mval,
 * Dptr, Dval, Lptr, Lval, Fval, res, rel, chd, tA, tB, uplo, side, diag)`
 * (src/Multifrontal.jl/src/divide.jl:417-435).  Every index array is passed
 * exactly as the reference stores it (Int64, 1-based, CSR pointers starting
 * at 1) and every value array in the reference's layout (Dval: per front an
 * nn x nn column-major block at Dptr[j]; Lval: an na x nn block at Lptr[j];
 * uplo = :L).  Paths in this header are relative to the CliqueTrees.jl tree.
 *
 * Error convention (cholesky.jl:384-388, utils.jl:775-785): functions return
 * MF_OK (0) or a negative status and never throw or abort; a numeric failure
 * is reported in *info as the first failing column in permuted order
 * (1-based), 0 on success — the value chol_impl! returns.  A plan and its
 * factors are not re-entrant: use one host thread per factor object.
 */
#ifndef MFHIP_H
#define MFHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mf_plan mf_plan;
typedef struct mf_factor mf_factor;

enum {
    MF_OK = 0,
    MF_EARG = -1,      /* invalid argument                                  */
    MF_EHIP = -2,      /* HIP runtime error (message in mf_last_error())    */
    MF_ENOMEM = -3,    /* device allocation failed                          */
    MF_ESTATE = -4     /* call out of order (e.g. solve before factorize)   */
};

enum { MF_FP64 = 0, MF_FP32 = 1 };          /* precision of the fronts       */
enum { MF_CHOLESKY = 0, MF_LDLT = 1 };      /* diag = :N / :U                */
enum { MF_REG_NONE = 0, MF_REG_DYNAMIC = 1, MF_REG_GMW81 = 2 };
enum { MF_UPLO_L = 0, MF_UPLO_U = 1 };     /* uplo = :L / :U                */

/* Message of the last failing call on this thread. */
const char* mf_last_error(void);

/*
 * Upload a ChordalSymbolic (src/Multifrontal.jl/src/chordal_symbolic.jl:15-28)
 * once: res = (resptr), sep = (sepptr, septgt), rel = (sepptr, reltgt),
 * chd = (chdptr [nfr+2], chdtgt [nfr]), Dptr, Lptr.  Builds the internal
 * front tree, the level schedule and all task lists; replaces the workspace
 * sizing of FactorizationWorkspace (cholesky.jl:15-24) and DivisionWorkspace
 * (divide.jl:13-18).  flags: reserved, pass 0.
 */
int mf_plan_create(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                   const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                   const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr,
                   int64_t flags, mf_plan** out);
int mf_plan_destroy(mf_plan* plan);

/*
 * mf_plan_create with planner parameters (no counterpart in the reference:
 * the autotuner's entry).  tuning[0..4]: panel steps per trailing-update
 * tile (1..64, default 8), parent rows per extend-add task (default 16384),
 * the front width from which update-matrix tiles are 128 x 128 (default
 * 128), workgroups assumed resident when ordering the tile-DAG's tickets
 * (default 512), explicit zeros allowed by a chain merge in parts per
 * million (default 50000), and tuning[5] a bit mask of tile-DAG task kinds
 * run at raised wave priority (bits 0..5: extend-add, small front, diagonal
 * tile, panel solve, trailing update, update matrix; bit 6: the look-ahead
 * trailing update of the panel chain; default 0; negative keeps it),
 * tuning[6] 1 to let each diagonal-tile task also run the panel solve of
 * the tile below it and the look-ahead update of the next diagonal tile, 0
 * (default) for separate tasks (negative keeps it), tuning[7] the resident
 * tile-DAG workgroups per CU (0, the default: as many as fit).  Entries
 * 0..4 <= 0 or beyond ntuning keep the default.  mf_plan_tuning reads back
 * the eight values and, as a ninth, the scheduler's estimate of the
 * tile-DAG launch in microseconds.
 */
int mf_plan_create_tuned(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                         const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                         const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr,
                         const int64_t* tuning, int64_t ntuning, mf_plan** out);
int mf_plan_tuning(const mf_plan* plan, int64_t* tuning, int64_t ntuning);

/* Plan statistics, in this order: n, nfr, internal fronts, fronts of the
 * small region, big fronts, big levels, small levels, max front size,
 * panel / U / W pool elements, GEMM tasks, extend-add tasks, diagonal-block
 * tasks, reference flops (sum nn^3/3 + na nn^2 + na^2 nn over the
 * reference's fronts), engine flops (the same over internal fronts), GEMM
 * flops, ndz, nlz, DAG tasks, reference flops of the fronts the tile-DAG
 * launch factors, reference flops of the small region, update-vector
 * elements per right-hand side of the solves.  nstats may be smaller than
 * the list. */
int mf_plan_stats(const mf_plan* plan, int64_t* stats, int64_t nstats);

/*
 * Device storage for one numeric factorization of a plan: the fields d,
 * Dval, Lval of ChordalFactorization (chordal_factorization.jl:1-21) plus the
 * engine's internal panels and update matrices.  precision: MF_FP64 or
 * MF_FP32 (fronts computed in FP32 with FP32 accumulation).
 */
int mf_factor_create(mf_plan* plan, int precision, mf_factor** out);
int mf_factor_destroy(mf_factor* f);
/* Device bytes the factor object holds now (its storage plus the solve,
 * refinement and staging workspaces allocated so far). */
int mf_factor_bytes(const mf_factor* f, int64_t* bytes);

/*
 * Storage triangle of the reference-layout arrays at this factor's boundary
 * (uplo of ChordalTriangular, cholesky.jl:373): MF_UPLO_L (default) keeps per
 * front the nn x nn lower block in Dval and the na x nn block below it in
 * Lval; MF_UPLO_U keeps their transposes, the nn x nn upper block and the
 * nn x na block to its right (chordal_triangular.jl:895-926, 1015-1055),
 * i.e. U = L'.  It applies to mf_factor_set_values, mf_factor_load,
 * mf_factor_get and the flat indices of mf_factor_set_pattern (register the
 * pattern after choosing the layout; a change drops the registered pattern).
 */
int mf_factor_set_uplo(mf_factor* f, int uplo);

/* All work of a factor object is ordered on this stream (a hipStream_t;
 * NULL restores the factor's own stream). */
int mf_factor_set_stream(mf_factor* f, void* stream);
int mf_factor_sync(mf_factor* f);
/* FP32 factor objects: solve in FP64 arithmetic (on = 1).  Each
 * factorization then also writes an FP64 copy of the factor (panels and
 * pivots widened, the 64 x 64 diagonal-tile inverses formed again in FP64),
 * which every later solve reads; the refinement's corrections then contract
 * faster than with FP32 solves.  No counterpart in the reference (its
 * factors have one element type).  Needs 8 bytes per factor entry more. */
int mf_factor_set_solve_fp64(mf_factor* f, int on);

/*
 * Values of P A P' in the reference layout, as left by copyto!(F, A)
 * (chordal_triangular.jl:852-893).  device != 0: the pointers are device
 * pointers (no PCIe transfer).
 */
int mf_factor_set_values(mf_factor* f, const double* Dval, const double* Lval, int device);

/*
 * A factor given by its values: the reference's ChordalFactorization built
 * from stored Dval/Lval/d (chordal_factorization.jl:1-21), as a saved factor
 * is restored.  Dval/Lval hold the factor in the reference layout (unit L
 * with the pivots in d for LDLt, diag = MF_LDLT); d may be NULL for
 * Cholesky (then d reads back as ones).  The diagonal blocks are inverted on
 * the device and the factor is ready for mf_solve / mf_solve_natural, as
 * after mf_factorize.  FP64 factor objects only.  device != 0: device
 * pointers, nothing waits on the host.
 */
int mf_factor_load(mf_factor* f, const double* Dval, const double* Lval, const double* d, int diag,
                   int device);

/*
 * Register flatindices(F, A) (chordal_factorization.jl:322-328) for a
 * pattern: flat[k] is the 1-based slot of nonzero k in [Dval; Lval] or 0.
 * mf_factor_refresh then does fill!(F, 0) followed by setflatindex! of every
 * nonzero (chordal_triangular.jl:837-849) on the device.
 */
int mf_factor_set_pattern(mf_factor* f, int64_t nnz, const int64_t* flat);
int mf_factor_refresh(mf_factor* f, const double* nzval, int device);

/*
 * Numeric factorization: cholesky!/ldlt! with NoPivot (cholesky.jl:80-101,
 * 181-203, 304-315), replacing chol_impl! (cholesky.jl:361-392).
 * diag: MF_CHOLESKY (:N) or MF_LDLT (:U).  signs: n pivot signs in permuted
 * order (NULL = Ones for Cholesky, Zeros for LDLt), host pointer.
 * reg_kind: MF_REG_NONE (NoRegularization, no_regularization.jl:27-42),
 * MF_REG_DYNAMIC (DynamicRegularization(delta, eps),
 * dynamic_regularization.jl:75-90) or MF_REG_GMW81 (GMW81(beta, delta),
 * regularization/gmw81.jl:37-155, with chol_factor!'s GMW81 form,
 * cholesky.jl:739-875): there delta is the minimum pivot and eps is beta; a
 * negative delta takes cbrt(eps(T)), a negative beta gmw81_beta of the
 * assembled matrix (computed on the device).  GMW81 with LDLt needs signs
 * (+-1, checksigns).  *info as chol_impl!.  The call returns once the work
 * is enqueued and *info is known (it synchronizes).
 */
int mf_factorize(mf_factor* f, int diag, const double* signs, int reg_kind, double delta,
                 double eps, int64_t* info);

/* Same, without synchronizing: *info is not produced; read it later with
 * mf_factor_info.  Used by the serving loop and by graph capture. */
int mf_factorize_async(mf_factor* f, int diag, const double* signs_dev, int reg_kind,
                       double delta, double eps);
int mf_factor_info(mf_factor* f, int64_t* info);

/* Copy the factor back in the reference layout (Dval, Lval, and for LDLt d).
 * Any pointer may be NULL.  device != 0: device pointers. */
int mf_factor_get(mf_factor* f, double* Dval, double* Lval, double* d, int device);

/*
 * Natural solve C <- L^-T [D^-1] L^-1 C (ldiv!(W, F::NaturalFactorization, C),
 * divide.jl:191-199, i.e. div_impl! forward, the diagonal solve and div_impl!
 * backward, divide.jl:417-450).  C is n x nrhs column-major with leading
 * dimension ldc, in permuted order.  device != 0: C is a device pointer.
 */
int mf_solve_natural(mf_factor* f, double* C, int64_t nrhs, int64_t ldc, int device);

/*
 * Full solve B <- A \ B = P' (L L')^-1 P B (ldiv!! divide.jl:1762-1764 with
 * mul!(C, P, B) multiply.jl:277-280); perm is F.perm (1-based, host).
 */
int mf_solve(mf_factor* f, const int64_t* perm, double* B, int64_t nrhs, int64_t ldb,
             int device);

/*
 * mf_solve without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  B is a
 * device pointer; perm as for mf_solve.  The permuted right-hand sides and
 * the update vectors live in a fixed device array of the library
 * (MF_NOALLOC_SCRATCH_BYTES per device, held until the process ends); a
 * solve that needs more returns MF_EARG.  The call only enqueues work on the
 * factor's stream and must be made with the factor's device current.  All
 * calls on one device share that array: calls that overlap on one GPU (other
 * streams, other factors, other host threads) race and must be serialized
 * by the caller.
 */
#define MF_NOALLOC_SCRATCH_BYTES (64ull << 20)
int mf_solve_noalloc(mf_factor* f, const int64_t* perm, double* B, int64_t nrhs, int64_t ldb);

/*
 * Iterative refinement (mixed precision: FP32 fronts, FP64 residual and
 * update; no counterpart in the reference, which factors in one precision).
 * mf_factor_set_matrix registers A (full symmetric CSC, 1-based, original
 * order, the same nonzero order as mf_factor_set_pattern so that
 * mf_factor_refresh also updates its values).  mf_refine: X <- A \ B by
 * X = F \ B, then X += F \ (B - A X) until max_j ||B_j - A X_j|| / ||B_j||
 * <= tol or maxit corrections; *iters = corrections done, *relres = final
 * worst relative residual.  device != 0: B and X are device pointers.
 */
int mf_factor_set_matrix(mf_factor* f, int64_t n, const int64_t* colptr, const int64_t* rowval,
                         const double* nzval, int device);
int mf_refine(mf_factor* f, const int64_t* perm, const double* B, int64_t ldb, double* X,
              int64_t ldx, int64_t nrhs, int device, double tol, int maxit, int* iters,
              double* relres);
/* Same for a batch of independent systems stacked in one factor: rows
 * [segptr[g], segptr[g+1]) (0-based, segptr[0] = 0, segptr[nseg] = n) are
 * system g, and the stopping test holds for every system and column. */
int mf_refine_segmented(mf_factor* f, const int64_t* perm, const double* B, int64_t ldb,
                        double* X, int64_t ldx, int64_t nrhs, int device, double tol, int maxit,
                        int64_t nseg, const int64_t* segptr, int* iters, double* relres);
/* Asynchronous form for serving loops (device pointers only, nothing waits
 * on the host): X = F \ B, then exactly `steps` corrections
 * X += F \ (B - A X), then the final residual.  nrm2 (device, 2 nseg nrhs
 * doubles) receives sum_i B(i,j)^2 at [j nseg + g] and sum_i R(i,j)^2 at
 * [nseg nrhs + j nseg + g] over the rows of segment g, so the caller reads
 * the relative residuals whenever it synchronizes the factor's stream.
 * segptr is host memory, as for mf_refine_segmented. */
int mf_refine_steps(mf_factor* f, const int64_t* perm, const double* B, int64_t ldb, double* X,
                    int64_t ldx, int64_t nrhs, int steps, int64_t nseg, const int64_t* segptr,
                    double* nrm2);

/*
 * Drop-in forms with chol_impl!'s own arguments (cholesky.jl:361-377):
 * Dval/Lval in (assembled A) and out (factor), d out (LDLt pivots; NULL is
 * accepted for Cholesky, whose d is Ones), signs in (NULL = the default),
 * uplo as for mf_factor_set_uplo, host memory, one call.  The plan keeps one
 * FP64 factor object for these calls (created on the first), so repeated
 * calls allocate nothing new; they are not re-entrant on one plan.
 */
int mf_chol_impl(mf_plan* plan, double* Dval, double* Lval, double* d, const double* signs,
                 int diag, int uplo, int reg_kind, double delta, double eps, int64_t* info);
/* div_impl! forward + (LDLt) diagonal + backward on host arrays
 * (divide.jl:417-435); d may be NULL for Cholesky. */
int mf_div_impl(mf_plan* plan, const double* Dval, const double* Lval, const double* d,
                double* C, int64_t nrhs, int64_t ldc, int diag, int uplo);

/* Selected inversion of a Cholesky factor: the entries of (P A P')^-1 on the
 * factor's pattern.  Replaces selinv_impl!(Mptr, Mval, Dptr, Dval, Lptr,
 * Lval, Fval, res, rel, chd, uplo) (src/Multifrontal.jl/src/selinv.jl:72-94),
 * which selinv!(F::ChordalCholesky) calls (selinv.jl:38-51); the reference
 * defines it for Cholesky factors only, and so does this.
 * mf_selinv: enqueues the inversion of the factor object's current factor on
 * its stream (the factor itself is kept); its work buffers are allocated on
 * the first call.  mf_selinv_get: the selected inverse in the reference's
 * layout of the factor object's uplo (Dval: the stored triangle of each
 * diagonal block, the other one 0; Lval: the off-diagonal blocks), host
 * arrays or, with device = 1, device pointers; synchronizes the stream.
 * mf_selinv_impl: the drop-in form: Dval/Lval hold a Cholesky factor in the
 * reference layout (uplo) and are overwritten with the selected inverse, as
 * selinv_impl! overwrites them; info = 0. */
int mf_selinv(mf_factor* f);
int mf_selinv_get(mf_factor* f, double* Dval, double* Lval, int device);
int mf_selinv_impl(mf_plan* plan, double* Dval, double* Lval, int uplo, int64_t* info);

/* Completion: Dval/Lval hold a matrix Sigma on the factor's pattern in the
 * reference layout (uplo; e.g. a selected inverse); they are overwritten
 * with the Cholesky factor of the inverse of Sigma's maximum-determinant
 * positive definite completion.  Replaces complete_impl!(Mptr, Mval, Dptr,
 * Dval, Lptr, Lval, Fval, res, rel, chd, uplo)
 * (src/Multifrontal.jl/src/complete.jl:46-72), which complete!(F)
 * (complete.jl:1-23) calls.  The fronts are independent given Sigma: each
 * front's reordered matrix [Sigma(sep, sep), Sigma21 J; J Sigma21', J
 * Sigma11 J] (J reverses the residual columns) is factored on the device as
 * one root of a forest, the big ones on the tile-DAG; the result is read
 * from the inverse of that factor.  info > 0: the first front in the
 * reference's order (root first) whose potrf! fails, with the reference's
 * local column (of Sigma(sep, sep), else of the reordered Schur complement);
 * Dval/Lval are then left unchanged.  Work buffers are kept in the plan
 * after the first call; calls on one plan must not overlap. */
int mf_complete_impl(mf_plan* plan, double* Dval, double* Lval, int uplo, int64_t* info);

/* Kernel timing (HIP events around each launch, on the factor's stream).
 * times[k] = total milliseconds, counts[k] = launches, flops[k] = algorithmic
 * work of kernel class k since the last reset (classes: MF_KC_*). */
enum { MF_KC_SMALL = 0, MF_KC_EA, MF_KC_DIAG, MF_KC_GEMM, MF_KC_COPYOUT, MF_KC_SOLVE,
       MF_KC_ASSEMBLE, MF_KC_DAG, MF_KC_SELINV, MF_KC_COUNT };
int mf_timing(mf_factor* f, int enable);
int mf_timing_read(mf_factor* f, double* times_ms, int64_t* counts, double* work, int reset);

/* Task-kind profile of the big-front tile-DAG kernel (no counterpart in the
 * reference; a diagnostic).  Kinds: extend-add, small front, diagonal tile,
 * panel-solve tile, trailing-update tile, update-matrix tile.  out[3k..3k+2]
 * = microseconds from the previous task's end to this task's start (ticket
 * selection and dependency waits), microseconds of work (summed over
 * workgroups), task count; out[18..20] = microseconds in the publishing
 * (release) step, in the acquire step and in the ticket fetch. */
enum { MF_DT_KINDS = 6 };
int mf_dag_profile(mf_factor* f, int enable);
int mf_dag_profile_read(mf_factor* f, double* out, int reset);
/* With the profile on, the last launch's per-ticket times: out[2t] = start
 * (wall-clock ticks) | workgroup << 48, out[2t+1] = end; n >= 2 x tasks.
 * mf_plan_dag: the tickets' tasks, 6 int32 each (kind, front, a, b, c, d). */
int mf_dag_trace_read(mf_factor* f, unsigned long long* out, int64_t n);
int mf_plan_dag(const mf_plan* plan, int32_t* tasks, int64_t n);
/* mf_plan_gemm: the tile products the DAG's panel-solve, trailing-update and
 * update-matrix tasks index (their field a), 4 int32 each: m, n, k, and
 * lower | wide << 1 | mode << 2 (mode 0: C -= A B', 1: C = A W'). */
int mf_plan_gemm(const mf_plan* plan, int32_t* out, int64_t n);
/* With the profile on, the last solve's per-ticket times (forward tickets,
 * then backward), as mf_dag_trace_read; n >= 2 x solve tasks.
 * mf_plan_solve_tasks: 5 int32 per ticket (kind 0 small front / 1 diagonal
 * block / 2 update rows / 3 up to four tiny fronts, front, a, b, c),
 * forward then backward.
 * mf_plan_fronts: 5 int64 per internal front (first permuted column, nn, na,
 * parent or -1, 1 if in the small region). */
int mf_solve_trace_read(mf_factor* f, unsigned long long* out, int64_t n);
int mf_plan_solve_tasks(const mf_plan* plan, int32_t* tasks, int64_t n);
int mf_plan_fronts(const mf_plan* plan, int64_t* out, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* MFHIP_H */
