// mf_plan.h — host-side plan of libmfhip.so: internal fronts, level
// schedule, memory layout and the static task lists of every launch.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mf_internal.h"

namespace mf {

// Planner parameters that shape the tile-DAG (chosen per pattern and
// precision by the autotuner; the defaults were tuned by hand on the 64^3
// 7-point Laplacian in FP64).
struct PlanTuning {
    int32_t upd_group = UPD_GROUP;      // panel steps per trailing-update tile (1..64)
    int32_t ea_rows = EA_ROWS;          // parent rows per extend-add task (>= EA_WIN)
    int32_t wide_k = WIDE_K;            // 128 x 128 update-matrix tiles when nn >= wide_k
    int32_t sim_workers = SIM_WORKERS;  // resident workgroups of the ticket simulation
    double zero_frac = 0.05;            // explicit zeros allowed by a chain merge
    int32_t prio = 0;                   // tile-DAG task kinds at raised wave priority (bit mask)
    int32_t fuse = 0;                   // diagonal tasks run the next panel solve and look-ahead update
    int32_t wg_per_cu = 0;              // resident tile-DAG workgroups per CU (0: as many as fit)
    bool factor_only = false;           // a plan for Cholesky/LDLt factorizations only (the
                                        // completion's forests): no solve task lists, no GMW81 order
};

// Task of the persistent triangular solves (one launch per direction).
enum : int32_t { PT_SMALL = 0, PT_DIAG = 1, PT_ROWS = 2, PT_TINY = 3, PT_SEPB = 4 };
struct PTask {
    int32_t kind;    // PT_SMALL: a whole small front; PT_DIAG: diagonal block rows
                     // [a, a+b) of a big front; PT_ROWS: update rows [a, b);
                     // PT_TINY: up to four fronts of nJ <= TINY_NJ (front, a, b,
                     // c; -1 when absent), one wavefront each; PT_SEPB (backward):
                     // the separator product of diagonal block a over the
                     // separator tiles [b, c), into the block's partial slot
    int32_t front;
    int32_t a, b;
    int32_t c;
};

struct Level {
    std::vector<int32_t> small_list;   // small fronts living in the big region
    std::vector<int32_t> big_list;
    int32_t smallL0 = 0, smallL1 = 0;  // range in level_small (device)
    int32_t bigL0 = 0, bigL1 = 0;      // range in level_big (device)
    int32_t ea0 = 0, ea1 = 0;          // EaTask range
};

struct HostPlan {
    int64_t n = 0, nfr = 0, ndz = 0, nlz = 0;
    // reference arrays, 0-based
    std::vector<int64_t> resptr, sepptr, septgt, Dptr, Lptr;
    std::vector<int32_t> pnt;          // reference front parent, -1 root
    // internal fronts
    std::vector<Front> fronts;
    std::vector<int32_t> rows, chd, rel;
    std::vector<int32_t> owner;        // reference front -> internal front
    std::vector<uint8_t> is_small;     // nJ <= SMALL_NJ
    int64_t panel_elems = 0, u_elems = 0, w_elems = 0, mv_elems1 = 0;
    // schedule
    std::vector<int32_t> small_order;  // small region, ticket order (children first)
    std::vector<int32_t> small_units;  // small_order in tickets: start << 3 | n tiny fronts (0: one front)
    std::vector<int32_t> level_small, level_big;
    std::vector<Level> levels;
    std::vector<EaTask> ea;
    std::vector<int32_t> ea_src;       // pairs (child, child column)
    std::vector<int32_t> wstart, wsoff;  // extend-add window tables per child (-1: none)
    std::vector<int32_t> sgstart, sgoff;  // solve gather tables per child of a big front (-1: none)
    std::vector<GemmTask> gemm;
    std::vector<DiagTask> diag;
    std::vector<PTask> ftask, btask;   // forward / backward solve tasks, ticket order
    std::vector<DagTask> dag;          // tile-DAG factorization of the big region
    std::vector<DagTask> dag_gmw;      // the same tasks in creation order (GMW81: a diagonal
                                       // tile waits for its whole block column)
    std::vector<DagFront> dagf;        // per internal front (small region: n_done = 1)
    int64_t dag_counters = 0;          // tile counters
    std::vector<int32_t> uwait_ptr, uwait;  // per front: fronts whose extend-add must end
                                            // before the front's U slot is written
    // dataflow form of the DAG: a task runs once its predecessors are done;
    // the extend-add (or small-front) tasks of a front are released together
    // by the front's gate (its children done, reused U slots consumed)
    std::vector<int32_t> dep0;              // per task: number of predecessor tasks
    std::vector<int32_t> succ_ptr, succ;    // successors (CSR over tasks)
    std::vector<int32_t> n_eau;             // per front: update-matrix extend-add groups
    std::vector<int32_t> eau_ptr, eau_wait; // per front: tiles waiting for all of them
    std::vector<int32_t> gate0;             // per front: children in the DAG + U-slot waits
    std::vector<int32_t> gtask0, gtask1;    // per front: the tasks its gate releases
    std::vector<int32_t> uwaiters_ptr, uwaiters;  // reverse of uwait
    std::vector<int8_t> hi;                 // per task: 1 = high-priority queue
    std::vector<int32_t> init_hi, init_lo;  // tasks ready at the start, per queue
    int32_t n_hi = 0, n_lo = 0;             // tasks per queue
    double sim_makespan_us = 0;             // list-scheduled length of the DAG (estimate)
    std::vector<int32_t> ntf, ntb;     // tasks per front (completion counts)
    std::vector<int64_t> mvoff1;       // per internal front: offset of Mv (na x 1)
    std::vector<int64_t> sboff1;       // per internal front: offset (per right-hand side) of its
                                       // backward separator partials (p x nch slots of 64 rows),
                                       // -1: the diagonal tasks compute the product themselves
    // copy-out
    std::vector<OutFront> out;
    std::vector<int32_t> septpos;
    PlanTuning tune;
    // statistics
    double ref_flops = 0, engine_flops = 0, gemm_flops = 0;
    int64_t n_small_region = 0, n_big = 0, max_nj = 0, small_levels = 0;
};

// Schedule of the selected inversion (k_selinv), built on first use.  Tasks
// in ticket order; a task waits only on counters that tasks with smaller
// tickets advance: its own front's earlier phases (every task of front J
// adds 1 to counter J when done) and its parent's completion.
struct SelPlan {
    std::vector<SelTask> task;
    std::vector<int32_t> total;            // per internal front: its task count
    std::vector<int64_t> xoff, ltoff, moff;  // per front (-1: none)
    int64_t x_elems = 0, lt_elems = 0, m_elems = 0;
    double flops = 0;                      // multiply-add flops of the task list (2 per FMA)
    int32_t max_depth = 0;
    int64_t n_gemm = 0, n_small = 0, n_gather = 0;
    int64_t ncounters = 0;                 // task counters the schedule uses (0: one per front)
};

// Builds the selected-inversion schedule of a plan (Cholesky factors).
std::string build_selinv(const HostPlan& H, SelPlan& S);

// Builds the completion's backward-solve schedule on a forest of roots:
// per big root J, pad_na[J] (a multiple of 64) padded separator rows and
// nnr[J] residual columns; the tasks run on k_selinv.
std::string build_cpl_solve(const HostPlan& H, const std::vector<int32_t>& pad_na,
                            const std::vector<int32_t>& nnr, SelPlan& S);

// Builds the plan; returns an empty string on success, else an error message.
std::string build_plan(HostPlan& P, int64_t n, int64_t nfr, const int64_t* resptr,
                       const int64_t* sepptr, const int64_t* septgt, const int64_t* reltgt,
                       const int64_t* chdptr, const int64_t* chdtgt, const int64_t* Dptr,
                       const int64_t* Lptr, const PlanTuning& tune);

// Tuning array of the C-ABI (mf_plan_create_tuned): missing or non-positive
// entries keep the defaults; returns an error message for out-of-range values.
std::string tuning_from(const int64_t* v, int64_t nv, PlanTuning* t);

}  // namespace mf
