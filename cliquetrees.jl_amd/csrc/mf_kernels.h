// mf_kernels.h — device argument blocks and host launch wrappers of the
// factorization and solve kernels (mf_factor.hip, mf_solve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mf_internal.h"
#include "mf_plan.h"

namespace mf {

// Device view of a plan (all arrays resident in HBM).
struct DevPlan {
    const Front* fronts;
    const int32_t* rows;
    const int32_t* chd;
    const int32_t* rel;
    const int32_t* small_order;
    const int32_t* small_units;  // tickets of the small region: start << 3 | tiny count (0: one front)
    const int32_t* level_small;
    const int32_t* level_big;
    const EaTask* ea;
    const int32_t* ea_src;
    const EaDst* ea_dst;        // per extend-add column
    const EaSrc* ea_srcx;       // per extend-add source (parallel to ea_src pairs)
    const int32_t* wstart;       // extend-add window tables
    const int32_t* wsoff;
    const int32_t* sgstart;      // solve gather tables: sgstart[sgoff[K] + b] = first row of
    const int32_t* sgoff;        // child K landing in or after row block b of its parent
    const GemmTask* gemm;
    const DiagTask* diag;
    const OutFront* out;
    const int32_t* septpos;
    const int64_t* mvoff1;
    const int64_t* sboff1;       // per front: backward separator partials (per right-hand side), -1: none
    const uint8_t* in_small;     // 1: front of the small region
    const int32_t* col2ref;      // global column -> reference front
    int32_t nJ;
    int32_t n;
    int32_t nfr;
    int32_t n_small_region;
};

// Device view of one factor object's storage.
template <typename T>
struct DevFactor {
    T* panel;
    T* U;
    T* W;
    T* d;                    // pivots (permuted order), LDLt
    const double* signs;     // NULL: default signs
    unsigned long long* info;  // min failing global column (1-based), ~0ull: none
    int32_t* flags;          // per internal front, persistent kernels
    uint32_t* ticket;
    int32_t* hang;           // set when a bounded wait gives up
    unsigned long long stall;  // wall-clock ticks without progress before a wait gives up
    int diag;                // 0 Cholesky, 1 LDLt
    int reg_kind;            // 0 none, 1 DynamicRegularization, 2 GMW81
    double delta, eps;       // GMW81: delta = minimum pivot (eps unused)
    const double* gmw;       // GMW81: gmw[2] = beta (written on the device before the factorization)
};

// Device view of the tile-DAG schedule of the big region (one factor object's
// counters; the task list and per-front constants belong to the plan).
struct DevDag {
    const DagTask* task;
    const DagFront* df;
    const int32_t* uwait_ptr;
    const int32_t* uwait;
    int32_t* cnt;      // tile counters
    int32_t* eacnt;    // finished extend-add tasks per front
    int32_t* eapcnt;   // finished extend-add tasks into panel columns per front
    unsigned long long* prof;  // NULL, or per task kind: wait, body (wall-clock ticks), count
    unsigned long long* trace; // NULL, or per ticket: start | workgroup << 48, end (wall clock)
    int32_t ntask;
    int32_t prio;      // bit k: tasks of kind k run at raised wave priority (bit DT_KINDS:
                       // the look-ahead trailing update of the panel chain)
    int32_t wg_per_cu; // resident workgroups per CU of the launch (0: as many as fit)
};

// ---- factorization (mf_factor.hip) ----
template <typename T>
void launch_fill(T* p, int64_t count, T value, hipStream_t s);
// uplo: 0 = :L, 1 = :U layout of Dval/Lval; adiag: assembled diagonal per
// permuted column (written by assembly and refresh, read by the copy-out).
template <typename T>
void launch_assemble_ref(const DevPlan& P, const DevFactor<T>& F, const double* Dval,
                         const double* Lval, int uplo, double* adiag, hipStream_t s);
template <typename T>
void launch_refresh(const DevFactor<T>& F, const int64_t* tgt_panel, const int64_t* tgt_diag,
                    const double* nzval, int64_t nnz, double* adiag, hipStream_t s);
template <typename T>
void launch_small(const DevPlan& P, const DevFactor<T>& F, const int32_t* list, const int32_t* units,
                  int count, hipStream_t s);
// invert_only: the block already holds the factor; only W is produced.
template <typename T>
void launch_diag(const DevPlan& P, const DevFactor<T>& F, int t0, int t1, bool invert_only,
                 hipStream_t s);
template <typename T>
void launch_dag(const DevPlan& P, const DevFactor<T>& F, const DevDag& D, hipStream_t s);
template <typename T>
void launch_copyout(const DevPlan& P, const DevFactor<T>& F, double* Dval, double* Lval,
                    double* dout, int uplo, const double* adiag, hipStream_t s);
void launch_widen(const float* src, double* dst, int64_t count, hipStream_t s);
// GMW81's beta (gmw81.jl:74-101) from the assembled panels into gmw[2]
// (gmw[0..1]: scratch of the reduction); n = matrix order, eps_t = eps(T).
template <typename T>
void launch_gmw_beta(const DevPlan& P, const DevFactor<T>& F, double* gmw, int64_t n, hipStream_t s);

// ---- solves (mf_solve.hip) ----
template <typename T>
struct DevSolve {
    T* B;        // n x nrhs, permuted order
    int64_t ldb;
    T* Mv;       // update vectors, na_J x nrhs at mvoff1[J]*nrhs
    int nrhs;
};
// Forward (children first) and backward (parents first) persistent solves,
// the D^-1 scaling in between for LDLt.
// trace: nullptr, or 2 (nf + nb) slots for per-ticket start / end ticks.
template <typename T>
void launch_solve(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, int nJ,
                  int32_t* done, int32_t* prog, const int32_t* ntf, const int32_t* ntb,
                  const PTask* ftask, int nf, const PTask* btask, int nb, int n,
                  unsigned long long* trace, hipStream_t s);
template <typename TB>
void launch_permute(const double* src, int64_t lds, TB* dst, int64_t ldd, const int64_t* perm,
                    int n, int nrhs, bool inverse, hipStream_t s);
void launch_permute_back(const void* src, int64_t lds, bool src_f32, double* dst, int64_t ldd,
                         const int64_t* perm, int n, int nrhs, hipStream_t s);

// ---- selected inversion (mf_selinv.hip) ----
// Device view of one factor object's selected-inversion state: the buffers
// of the schedule (SP_PANEL and SP_W are the factor's own), its task list
// and per-front offsets (the plan's), the per-front task counters.
template <typename T>
struct DevSel {
    T* base[SP_N];
    const SelTask* task;
    const int64_t* xoff;     // per front: X (and T) offset, -1 for small fronts
    const int64_t* moff;     // per front: M offset, -1 when na = 0
    int32_t* cnt;            // finished tasks per front
    uint32_t* ticket;
    int32_t* hang;           // the watchdog record (mf_sync.h)
    unsigned long long stall;
    int32_t ntask;
};
template <typename T>
void launch_selinv(const DevPlan& P, const DevSel<T>& S, hipStream_t s);

// ---- completion (mf_complete.hip), FP64 ----
// K of every reference front (the reordered front matrix whose Cholesky
// factor carries the completion) from Sigma in the reference layout.
void launch_cpl_gather(const DevPlan& P, int m, const int32_t* fr, const double* Dval, const double* Lval,
                       const int32_t* septgt, const int64_t* koff, double* K, hipStream_t s);
void launch_cpl_extract(const DevPlan& P, int m, const int32_t* fr, const double* X, const int64_t* xoff,
                        const double* L, const int64_t* loff, double* Dval, double* Lval, hipStream_t s);

// ---- iterative refinement (mf_refine.hip), FP64 ----
void launch_residual(int64_t n, const int64_t* colptr, const int64_t* rowval, const double* val,
                     const double* X, int64_t ldx, const double* B, int64_t ldb, double* R,
                     int64_t ldr, int nrhs, hipStream_t s);
void launch_colnorm2(const int64_t* seg, int nseg, const double* M, int64_t ld, int nrhs,
                     double* out, hipStream_t s);
void launch_axpy(int64_t n, const double* D, int64_t ldd, double* X, int64_t ldx, int nrhs,
                 hipStream_t s);
void launch_copy_cols(int64_t n, const double* S, int64_t lds, double* D, int64_t ldd, int nrhs,
                      hipStream_t s);

}  // namespace mf
