// mf_api.hip — the extern "C" boundary of libmfhip.so (include/mfhip.h):
// plan upload, factor storage, the launch sequences of the factorization
// and of the solves, kernel timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mfhip.h"
#include "mf_kernels.h"
#include "mf_plan.h"

using namespace mf;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess)                                                        \
            return fail(MF_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));    \
    } while (0)

template <typename V>
int upload(const std::vector<V>& h, V** d, std::vector<void*>& allocs) {
    size_t bytes = std::max<size_t>(h.size() * sizeof(V), 16);
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return fail(MF_ENOMEM, "hipMalloc failed (plan)");
    allocs.push_back(p);
    if (!h.empty() && hipMemcpy(p, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice) != hipSuccess)
        return fail(MF_EHIP, "hipMemcpy failed (plan)");
    *d = (V*)p;
    return MF_OK;
}

}  // namespace

// Device scratch of mf_solve_noalloc: the permuted right-hand sides and the
// solves' update vectors, for callers that cannot allocate device memory.
// Every device holds its own copy for the life of the process.  Host code
// never uses this name as an address: it takes the device address once per
// device with hipGetSymbolAddress (scratch_address below).
static __device__ __align__(256) unsigned char scratch[MF_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (taken on the first call
// per device, with that device current).
int scratch_address(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(MF_EARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MF_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MF_OK;
}

}  // namespace

struct mf_plan {
    HostPlan H;
    DevPlan D{};
    std::vector<void*> allocs;
    std::vector<uint8_t> in_small;
    std::vector<int32_t> col2ref;
    PTask* d_ftask = nullptr;
    PTask* d_btask = nullptr;
    int32_t* d_ntf = nullptr;
    int32_t* d_ntb = nullptr;
    DagTask* d_dag = nullptr;
    DagTask* d_dag_gmw = nullptr;  // the same tasks in creation order (GMW81 factorizations)
    DagFront* d_dagf = nullptr;
    int32_t* d_uwait_ptr = nullptr;
    int32_t* d_uwait = nullptr;
    double ea_bytes = 0, small_flops = 0, diag_flops = 0;
    // reference flops (sum nn^3/3 + na nn^2 + na^2 nn over the reference's own
    // fronts) of the fronts each launch factors: the small region, the DAG
    double ref_flops_small = 0, ref_flops_dag = 0;
    mf_factor* dropin = nullptr;   // FP64 factor object reused by mf_chol_impl / mf_div_impl
    // selected inversion: schedule built and uploaded on first use
    SelPlan* sel = nullptr;
    struct CplState* cpl = nullptr;  // completion: its forest of front matrices (first use)
    SelTask* d_sel = nullptr;
    int64_t* d_selx = nullptr;
    int64_t* d_selm = nullptr;
    ~mf_plan();
};

struct TimedLaunch {
    int kc;
    hipEvent_t a, b;
};

struct mf_factor {
    mf_plan* plan = nullptr;
    int precision = MF_FP64;
    int device = 0;             // the device current at mf_factor_create
    int uplo = MF_UPLO_L;       // layout of Dval/Lval at the boundary (:L or :U)
    unsigned long long stall = 0;  // wall-clock ticks of the bounded waits' watchdog (10 s)
    double* adiag = nullptr;    // assembled diagonal per permuted column
    hipStream_t own = nullptr, stream = nullptr;
    void* panel = nullptr;
    void* U = nullptr;
    void* W = nullptr;
    void* d = nullptr;          // pivots in the front precision
    double* Dref = nullptr;     // reference-layout Dval / Lval / d (double)
    double* Lref = nullptr;
    double* dref = nullptr;
    double* signs = nullptr;
    unsigned long long* info = nullptr;
    int32_t* flags = nullptr;
    int32_t* prog = nullptr;
    uint32_t* ticket = nullptr;
    int32_t* hang = nullptr;
    double* gmw = nullptr;      // GMW81: reduction scratch, then beta (gmw[2])
    int32_t* cnt = nullptr;     // tile-DAG counters
    int32_t* eacnt = nullptr;   // 2 (fronts + 1): all extend-add tasks, then panel ones
    unsigned long long* prof = nullptr;  // task-kind profile of the tile-DAG (NULL: off)
    // FP64 matrix of the refinement residual (full symmetric CSC, original order)
    int64_t A_n = -1, A_nnz = 0;
    int64_t* A_colptr = nullptr;
    int64_t* A_rowval = nullptr;
    double* A_val = nullptr;
    void* ir_ws = nullptr;     // B copy, X, R (n x nrhs each) + norms
    size_t ir_ws_bytes = 0;
    unsigned long long* prof_buf = nullptr;
    unsigned long long* trace_buf = nullptr;   // per DAG ticket: start, end
    unsigned long long* strace_buf = nullptr;  // per solve ticket (forward, then backward)
    int64_t* tgt_panel = nullptr;
    int64_t* tgt_diag = nullptr;   // per nonzero: its permuted column if diagonal, else -1
    int64_t nnz = -1;
    double* stage = nullptr;    // staging for host nzval / values
    size_t stage_bytes = 0;
    void* Mv = nullptr;
    size_t Mv_bytes = 0;
    void* Bws = nullptr;
    size_t Bws_bytes = 0;
    bool solve64 = false;       // FP32 factor: solves in FP64 arithmetic on a widened copy
    void* sh_panel = nullptr;   // that copy: panels, diagonal-tile inverses, pivots (FP64)
    void* sh_W = nullptr;
    void* sh_d = nullptr;
    bool sh_valid = false;
    void* sel_pool = nullptr;   // selected inversion: X, T, Lt, M, Sigma panels
    size_t sel_bytes = 0;
    int32_t* sel_cnt = nullptr; // finished selected-inversion tasks per front
    bool sel_valid = false;     // sel_pool holds the selected inverse of the current factor
    int64_t* perm_dev = nullptr;
    std::vector<int64_t> perm_host;   // the permutation held in perm_dev
    int diag = MF_CHOLESKY;
    int reg_kind = 0;
    double delta = 0, eps = 0;
    bool factored = false;
    bool ref_stale = false;    // Dref/Lref/dref not yet copied out of the panels
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> evpool;
    double t_ms[MF_KC_COUNT] = {0};
    int64_t t_cnt[MF_KC_COUNT] = {0};
    double t_work[MF_KC_COUNT] = {0};
    std::vector<void*> allocs;
    size_t alloc_bytes = 0;     // bytes of the allocations in allocs
    ~mf_factor() {
        for (auto& t : pending) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        for (auto e : evpool) (void)hipEventDestroy(e);
        for (void* p : allocs) (void)hipFree(p);
        if (Mv) (void)hipFree(Mv);
        if (ir_ws) (void)hipFree(ir_ws);
        if (A_colptr) (void)hipFree(A_colptr);
        if (A_rowval) (void)hipFree(A_rowval);
        if (A_val) (void)hipFree(A_val);
        if (Bws) (void)hipFree(Bws);
        if (stage) (void)hipFree(stage);
        if (tgt_panel) (void)hipFree(tgt_panel);
        if (tgt_diag) (void)hipFree(tgt_diag);
        if (sel_pool) (void)hipFree(sel_pool);
        if (sh_panel) (void)hipFree(sh_panel);
        if (sh_W) (void)hipFree(sh_W);
        if (sh_d) (void)hipFree(sh_d);
        if (sel_cnt) (void)hipFree(sel_cnt);
        if (own) (void)hipStreamDestroy(own);
    }
};

// Device memory handed out from one buffer, 256-byte aligned.
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
};

// The completion's forest (complete_impl!, mf_complete.hip): per reference
// front j one root front holding its reordered front matrix K (nJ' x nJ',
// the separator part of a big front padded to a multiple of 64 rows), in
// the reference's order of the completion (root first), cut into chunks
// whose K fit a memory budget.  A chunk keeps its plan, the original front
// of each of its roots, per root the panel offset of K (factored in place)
// and the offset of its backward solves Y (-1: small front, inverted by
// k_cpl_extract), and the schedule of those solves.
struct CplChunk {
    mf_plan* syn = nullptr;
    int64_t t0 = 0, t1 = 0;       // root-first positions [t0, t1)
    std::vector<int64_t> resptr;  // the chunk's column ranges (1-based), for info
    std::vector<int64_t> sep;     // per root: na and the padded na'
    std::vector<int64_t> sepp;
    int32_t* d_fr = nullptr;
    int64_t* d_poff = nullptr;
    int64_t* d_xoff = nullptr;
    SelTask* d_task = nullptr;    // the backward solves (k_selinv tasks)
    int32_t ntask = 0;
    int64_t nctr = 0;             // their counters
    int64_t y_elems = 0;          // Y (and the same again for U and for R)
};

struct CplState {
    std::vector<CplChunk> chunks;
    std::vector<void*> allocs;
    Arena arena;  // one chunk's factor object, Y, U, R and the solve counters
    double* d_eye = nullptr;  // a 64 x 64 identity
    int32_t* d_septgt = nullptr;
    double* d_sig = nullptr;  // Sigma (ndz + nlz)
    double* d_out = nullptr;  // the result (ndz + nlz)
    ~CplState() {
        for (CplChunk& c : chunks)
            if (c.syn) (void)mf_plan_destroy(c.syn);
        for (void* p : allocs) (void)hipFree(p);
    }
};

mf_plan::~mf_plan() {
    delete dropin;
    delete sel;
    delete cpl;
    for (void* p : allocs) (void)hipFree(p);
}

namespace {

hipEvent_t get_event(mf_factor* f) {
    if (!f->evpool.empty()) {
        hipEvent_t e = f->evpool.back();
        f->evpool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

// Times one launch (class kc) when timing is on; nothing is launched or
// recorded for an empty task range.
template <typename Fn>
void timed(mf_factor* f, int kc, double work, int64_t count, Fn&& fn) {
    if (count <= 0) return;
    if (!f->timing) {
        fn();
        return;
    }
    TimedLaunch t{kc, get_event(f), get_event(f)};
    (void)hipEventRecord(t.a, f->stream);
    fn();
    (void)hipEventRecord(t.b, f->stream);
    f->pending.push_back(t);
    f->t_work[kc] += work;
}

void drain_timing(mf_factor* f) {
    if (f->pending.empty()) return;
    (void)hipStreamSynchronize(f->stream);
    for (auto& t : f->pending) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        f->t_ms[t.kc] += ms;
        f->t_cnt[t.kc] += 1;
        f->evpool.push_back(t.a);
        f->evpool.push_back(t.b);
    }
    f->pending.clear();
}

template <typename T>
DevFactor<T> dev_factor(mf_factor* f) {
    DevFactor<T> F{};
    F.panel = (T*)f->panel;
    F.U = (T*)f->U;
    F.W = (T*)f->W;
    F.d = (T*)f->d;
    F.signs = f->signs;
    F.info = f->info;
    F.flags = f->flags;
    F.ticket = f->ticket;
    F.hang = f->hang;
    F.stall = f->stall;
    F.diag = f->diag;
    F.reg_kind = f->reg_kind;
    F.delta = f->delta;
    F.eps = f->eps;
    F.gmw = f->gmw;
    return F;
}

// An FP32 factor object with FP64 solves: the solves read the widened copy.
bool solves64(const mf_factor* f) { return f->precision == MF_FP32 && f->solve64; }

DevFactor<double> dev_factor_solve64(mf_factor* f) {
    DevFactor<double> F = dev_factor<double>(f);
    F.panel = (double*)f->sh_panel;
    F.W = (double*)f->sh_W;
    F.d = (double*)f->sh_d;
    return F;
}

int ensure(void** p, size_t* have, size_t need) {
    if (*have >= need && *p) return MF_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    if (hipMalloc(p, std::max<size_t>(need, 16)) != hipSuccess) return fail(MF_ENOMEM, "hipMalloc failed (workspace)");
    *have = std::max<size_t>(need, 16);
    return MF_OK;
}

// The FP64 copy of an FP32 factor for FP64-arithmetic solves: panels and
// pivots widened, the 64 x 64 diagonal-tile inverses formed again in FP64
// from the widened panels (buffers allocated by mf_factor_set_solve_fp64).
void enqueue_widen(mf_factor* f) {
    const HostPlan& H = f->plan->H;
    hipStream_t s = f->stream;
    launch_widen((const float*)f->panel, (double*)f->sh_panel, H.panel_elems, s);
    launch_widen((const float*)f->d, (double*)f->sh_d, H.n, s);
    launch_diag<double>(f->plan->D, dev_factor_solve64(f), 0, (int)H.diag.size(), true, s);
    f->sh_valid = true;
}

// Launch sequence of one numeric factorization (values already in the panels).
template <typename T>
void enqueue_factor(mf_factor* f) {
    mf_plan* pl = f->plan;
    const HostPlan& H = pl->H;
    const DevPlan& P = pl->D;
    hipStream_t s = f->stream;
    DevFactor<T> F = dev_factor<T>(f);
    launch_fill<unsigned long long>(f->info, 1, ~0ull, s);
    launch_fill<int32_t>(f->hang, 16, 0, s);  // the watchdog record (the solves keep it)
    launch_fill<int32_t>(f->flags, std::max<int>(P.nJ, 1), 0, s);
    launch_fill<uint32_t>(f->ticket, 1, 0u, s);
    if (f->reg_kind == MF_REG_GMW81) {
        // initialize(L, signs, GMW81()) (gmw81.jl:50-63): beta of the assembled matrix
        if (f->eps < 0) launch_gmw_beta<T>(P, F, f->gmw, H.n, s);
        else launch_fill<double>(f->gmw + 2, 1, f->eps, s);
    }
    timed(f, MF_KC_SMALL, pl->ref_flops_small, H.n_small_region, [&] {
        launch_small<T>(P, F, P.small_order, P.small_units, (int)H.small_units.size(), s);
    });
    if (!H.dag.empty()) {
        launch_fill<int32_t>(f->cnt, H.dag_counters, 0, s);
        const size_t nfr1 = H.fronts.size() + 1;
        launch_fill<int32_t>(f->eacnt, 2 * nfr1, 0, s);
        launch_fill<uint32_t>(f->ticket, 1, 0u, s);
        // GMW81: the ticket order where a diagonal tile follows every update of its block column
        DevDag D{f->reg_kind == MF_REG_GMW81 ? pl->d_dag_gmw : pl->d_dag, pl->d_dagf, pl->d_uwait_ptr, pl->d_uwait, f->cnt, f->eacnt, f->eacnt + nfr1,
                 f->prof, f->prof ? f->trace_buf : nullptr,
                 (int32_t)H.dag.size(), H.tune.prio, H.tune.wg_per_cu};
        timed(f, MF_KC_DAG, pl->ref_flops_dag, 1, [&] { launch_dag<T>(P, F, D, s); });
    }
    if constexpr (sizeof(T) == 4) {
        if (f->solve64) enqueue_widen(f);
    }
}

// The reference-layout Dval/Lval/d are written from the panels only when a
// caller reads them (the solves and the refinement use the panels), so a
// factor + solve step moves no bytes for them.
template <typename T>
void enqueue_copyout(mf_factor* f) {
    const HostPlan& H = f->plan->H;
    DevFactor<T> F = dev_factor<T>(f);
    timed(f, MF_KC_COPYOUT, 8.0 * 2.0 * (double)(H.ndz + H.nlz), H.nfr, [&] {
        launch_copyout<T>(f->plan->D, F, f->Dref, f->Lref, f->dref, f->uplo, f->adiag, f->stream);
    });
}

void sync_ref(mf_factor* f) {
    if (!f->ref_stale) return;
    if (f->precision == MF_FP32) enqueue_copyout<float>(f);
    else enqueue_copyout<double>(f);
    f->ref_stale = false;
}

// mv: the update-vector workspace (mv_elems1 x nrhs elements); NULL takes
// the factor's own.
template <typename T>
void enqueue_solve(mf_factor* f, T* B, int64_t ldb, int nrhs, void* mv = nullptr) {
    mf_plan* pl = f->plan;
    const HostPlan& H = pl->H;
    const DevPlan& P = pl->D;
    DevFactor<T> F = dev_factor<T>(f);
    if constexpr (sizeof(T) == 8) {
        if (solves64(f)) F = dev_factor_solve64(f);
    }
    DevSolve<T> S{B, ldb, (T*)(mv ? mv : f->Mv), nrhs};
    double bytes = (double)sizeof(T) * (2.0 * (double)(H.ndz + H.nlz) + 4.0 * H.n * nrhs);
    timed(f, MF_KC_SOLVE, bytes, 1, [&] {
        launch_solve<T>(P, F, S, P.nJ, f->flags, f->prog, pl->d_ntf, pl->d_ntb, pl->d_ftask,
                        (int)H.ftask.size(), pl->d_btask, (int)H.btask.size(), (int)H.n,
                        f->prof ? f->strace_buf : nullptr, f->stream);
    });
}

size_t elt(const mf_factor* f) { return f->precision == MF_FP32 ? 4 : 8; }
// element size of the solves' workspaces
size_t solve_elt(const mf_factor* f) { return f->precision == MF_FP32 && !solves64(f) ? 4 : 8; }

// The watchdog record of mf_sync.h's give_up, with the awaited counter named.
std::string hang_message(const mf_factor* f, const int32_t* h) {
    const unsigned long long a = (unsigned long long)(uint32_t)h[4] | ((unsigned long long)(uint32_t)h[5] << 32);
    const HostPlan& H = f->plan->H;
    const size_t nf1 = H.fronts.size() + 1;
    struct Buf {
        const char* name;
        const void* base;
        size_t n;
    } bufs[] = {{"front flag/done", f->flags, nf1},
                {"solve progress", f->prog, nf1},
                {"tile counter", f->cnt, (size_t)H.dag_counters + 1},
                {"extend-add counter", f->eacnt, 2 * nf1}};
    std::string where = "an unknown counter";
    for (const Buf& b : bufs) {
        const unsigned long long b0 = (unsigned long long)b.base;
        if (a >= b0 && a < b0 + 4 * b.n) {
            where = std::string(b.name) + " [" + std::to_string((a - b0) / 4) + "]";
            break;
        }
    }
    return "a bounded device wait gave up: " + where + " = " + std::to_string(h[2]) + ", awaited >= " +
           std::to_string(h[1]) + (h[3] >= 0 ? " (ticket " + std::to_string(h[3]) + ")" : std::string());
}

int read_info(mf_factor* f, int64_t* info) {
    unsigned long long v = 0;
    int32_t hang[6] = {0, 0, 0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(&v, f->info, sizeof(v), hipMemcpyDeviceToHost, f->stream));
    HIPCHK(hipMemcpyAsync(hang, f->hang, sizeof(hang), hipMemcpyDeviceToHost, f->stream));
    HIPCHK(hipStreamSynchronize(f->stream));
    if (hang[0]) return fail(MF_EHIP, hang_message(f, hang));
    if (info) *info = (v == ~0ull) ? 0 : (int64_t)v;
    return MF_OK;
}

}  // namespace

extern "C" {

const char* mf_last_error(void) { return g_err.c_str(); }

static int plan_create(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                       const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                       const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr,
                       const PlanTuning& tune, mf_plan** out) {
    if (!out || !resptr || !sepptr || !chdptr || !Dptr || !Lptr || (nfr > 0 && !chdtgt))
        return fail(MF_EARG, "null argument");
    *out = nullptr;
    mf_plan* pl = new mf_plan();
    std::string err = build_plan(pl->H, n, nfr, resptr, sepptr, septgt, reltgt, chdptr, chdtgt,
                                 Dptr, Lptr, tune);
    if (!err.empty()) {
        delete pl;
        return fail(MF_EARG, "plan: " + err);
    }
    HostPlan& H = pl->H;
    int rc = MF_OK;
    DevPlan& D = pl->D;
    pl->in_small.assign(H.fronts.size(), 0);
    pl->col2ref.assign(H.n, 0);
    for (int64_t j = 0; j < H.nfr; ++j)
        for (int64_t g = H.resptr[j]; g < H.resptr[j + 1]; ++g) pl->col2ref[g] = (int32_t)j;
    for (int32_t J : H.small_order) pl->in_small[J] = 1;
    std::vector<int32_t> ea_src = H.ea_src;
    std::vector<EaDst> ea_dst(H.ea.size());
    for (size_t e = 0; e < H.ea.size(); ++e) {
        const Front& F = H.fronts[H.ea[e].front];
        const int32_t pc = H.ea[e].col;
        const bool to_u = pc >= F.nn;
        ea_dst[e].off = to_u ? F.uoff + (int64_t)(pc - F.nn) * F.na - F.nn
                             : F.poff + (int64_t)pc * (F.nn + F.na);
        ea_dst[e].to_u = to_u ? 1 : 0;
        ea_dst[e].pad = 0;
    }
    std::vector<EaSrc> ea_srcx(H.ea_src.size() / 2);
    for (const EaTask& e : H.ea) {
        const int32_t w = e.row0 / EA_WIN;
        for (int32_t q = e.src0; q < e.src1; ++q) {
            const int32_t K = H.ea_src[2 * q], tcol = H.ea_src[2 * q + 1];
            const Front& Kf = H.fronts[K];
            EaSrc& x = ea_srcx[q];
            x.u_off = Kf.uoff + (int64_t)tcol * Kf.na;
            x.rel_off = Kf.reloff;
            x.ws_off = H.wsoff[K];
            x.tcol = tcol;
            x.i0 = std::max(H.wstart[H.wsoff[K] + w], tcol);
            x.i1 = H.wstart[H.wsoff[K] + w + 1];
            x.pad = 0;
        }
    }
    if ((rc = upload(H.fronts, (Front**)&D.fronts, pl->allocs)) ||
        (rc = upload(ea_dst, (EaDst**)&D.ea_dst, pl->allocs)) ||
        (rc = upload(ea_srcx, (EaSrc**)&D.ea_srcx, pl->allocs)) ||
        (rc = upload(H.rows, (int32_t**)&D.rows, pl->allocs)) ||
        (rc = upload(H.chd, (int32_t**)&D.chd, pl->allocs)) ||
        (rc = upload(H.rel, (int32_t**)&D.rel, pl->allocs)) ||
        (rc = upload(H.small_order, (int32_t**)&D.small_order, pl->allocs)) ||
        (rc = upload(H.small_units, (int32_t**)&D.small_units, pl->allocs)) ||
        (rc = upload(H.level_small, (int32_t**)&D.level_small, pl->allocs)) ||
        (rc = upload(H.level_big, (int32_t**)&D.level_big, pl->allocs)) ||
        (rc = upload(H.ea, (EaTask**)&D.ea, pl->allocs)) ||
        (rc = upload(ea_src, (int32_t**)&D.ea_src, pl->allocs)) ||
        (rc = upload(H.wstart, (int32_t**)&D.wstart, pl->allocs)) ||
        (rc = upload(H.wsoff, (int32_t**)&D.wsoff, pl->allocs)) ||
        (rc = upload(H.sgstart, (int32_t**)&D.sgstart, pl->allocs)) ||
        (rc = upload(H.sgoff, (int32_t**)&D.sgoff, pl->allocs)) ||
        (rc = upload(H.gemm, (GemmTask**)&D.gemm, pl->allocs)) ||
        (rc = upload(H.diag, (DiagTask**)&D.diag, pl->allocs)) ||
        (rc = upload(H.ftask, &pl->d_ftask, pl->allocs)) ||
        (rc = upload(H.btask, &pl->d_btask, pl->allocs)) ||
        (rc = upload(H.ntf, &pl->d_ntf, pl->allocs)) ||
        (rc = upload(H.ntb, &pl->d_ntb, pl->allocs)) ||
        (rc = upload(H.dag, &pl->d_dag, pl->allocs)) ||
        (rc = upload(H.dag_gmw, &pl->d_dag_gmw, pl->allocs)) ||
        (rc = upload(H.dagf, &pl->d_dagf, pl->allocs)) ||
        (rc = upload(H.uwait_ptr, &pl->d_uwait_ptr, pl->allocs)) ||
        (rc = upload(H.uwait, &pl->d_uwait, pl->allocs)) ||
        (rc = upload(H.out, (OutFront**)&D.out, pl->allocs)) ||
        (rc = upload(H.septpos, (int32_t**)&D.septpos, pl->allocs)) ||
        (rc = upload(H.mvoff1, (int64_t**)&D.mvoff1, pl->allocs)) ||
        (rc = upload(H.sboff1, (int64_t**)&D.sboff1, pl->allocs)) ||
        (rc = upload(pl->in_small, (uint8_t**)&D.in_small, pl->allocs)) ||
        (rc = upload(pl->col2ref, (int32_t**)&D.col2ref, pl->allocs))) {
        delete pl;
        return rc;
    }
    D.nJ = (int32_t)H.fronts.size();
    D.n = (int32_t)H.n;
    D.nfr = (int32_t)H.nfr;
    D.n_small_region = (int32_t)H.n_small_region;
    // algorithmic work per kernel class
    for (size_t J = 0; J < H.fronts.size(); ++J) {
        const Front& F = H.fronts[J];
        double a = F.nn, b = F.na;
        if (H.is_small[J]) pl->small_flops += a * a * a / 3.0 + b * a * a + b * b * a;
        else
            for (int32_t q = F.chd0; q < F.chd1; ++q) {
                double m = H.fronts[H.chd[q]].na;
                pl->ea_bytes += 3.0 * 8.0 * m * (m + 1) / 2.0 + 4.0 * m;
            }
    }
    for (const DiagTask& t : H.diag) pl->diag_flops += (double)t.kb * t.kb * t.kb / 3.0;
    for (const OutFront& o : H.out) {
        const double a = o.nn, b = o.na, w = a * a * a / 3.0 + b * a * a + b * b * a;
        if (pl->in_small[o.J]) pl->ref_flops_small += w;
        else pl->ref_flops_dag += w;
    }
    *out = pl;
    return MF_OK;
}

int mf_plan_create(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                   const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                   const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr,
                   int64_t flags, mf_plan** out) {
    (void)flags;
    return plan_create(n, nfr, resptr, sepptr, septgt, reltgt, chdptr, chdtgt, Dptr, Lptr,
                       PlanTuning(), out);
}

int mf_plan_create_tuned(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                         const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                         const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr,
                         const int64_t* tuning, int64_t ntuning, mf_plan** out) {
    PlanTuning t;
    std::string err = tuning_from(tuning, ntuning, &t);
    if (!err.empty()) return fail(MF_EARG, "tuning: " + err);
    return plan_create(n, nfr, resptr, sepptr, septgt, reltgt, chdptr, chdtgt, Dptr, Lptr, t, out);
}

int mf_plan_tuning(const mf_plan* pl, int64_t* tuning, int64_t ntuning) {
    if (!pl || !tuning) return fail(MF_EARG, "null argument");
    const PlanTuning& t = pl->H.tune;
    const int64_t v[] = {t.upd_group, t.ea_rows, t.wide_k, t.sim_workers,
                         (int64_t)llround(t.zero_frac * 1e6), t.prio, t.fuse, t.wg_per_cu,
                         (int64_t)llround(pl->H.sim_makespan_us)};
    for (int64_t i = 0; i < ntuning && i < (int64_t)(sizeof(v) / sizeof(v[0])); ++i) tuning[i] = v[i];
    return MF_OK;
}

int mf_plan_destroy(mf_plan* plan) {
    delete plan;
    return MF_OK;
}

int mf_plan_stats(const mf_plan* pl, int64_t* st, int64_t nst) {
    if (!pl || !st) return fail(MF_EARG, "null argument");
    const HostPlan& H = pl->H;
    int64_t v[] = {H.n, H.nfr, (int64_t)H.fronts.size(), H.n_small_region, H.n_big,
                   (int64_t)H.levels.size(), H.small_levels, H.max_nj, H.panel_elems, H.u_elems,
                   H.w_elems, (int64_t)H.gemm.size(), (int64_t)H.ea.size(),
                   (int64_t)H.diag.size(), (int64_t)llround(H.ref_flops),
                   (int64_t)llround(H.engine_flops), (int64_t)llround(H.gemm_flops), H.ndz,
                   H.nlz, (int64_t)H.dag.size(), (int64_t)llround(pl->ref_flops_dag),
                   (int64_t)llround(pl->ref_flops_small), H.mv_elems1};
    int64_t m = (int64_t)(sizeof(v) / sizeof(v[0]));
    for (int64_t i = 0; i < nst && i < m; ++i) st[i] = v[i];
    return MF_OK;
}

}  // extern "C"

// The factor object's device buffers in allocation order; refs: the
// reference-layout copies Dref/Lref (a factor carved from an arena, used
// only inside the library, has none).
static std::vector<std::pair<void**, size_t>> factor_buffers(mf_factor* f, const HostPlan& H, size_t es,
                                                            bool refs) {
    const size_t nf = H.fronts.size() + 1, n1 = (size_t)(H.n + 1);
    std::vector<std::pair<void**, size_t>> v = {
        {&f->panel, es * (size_t)H.panel_elems}, {&f->U, es * (size_t)H.u_elems},
        {&f->W, es * (size_t)H.w_elems},         {&f->d, es * n1}};
    if (refs) {
        v.push_back({(void**)&f->Dref, 8 * (size_t)H.ndz});
        v.push_back({(void**)&f->Lref, 8 * (size_t)H.nlz});
    }
    const std::pair<void**, size_t> rest[] = {
        {(void**)&f->dref, 8 * n1},   {(void**)&f->adiag, 8 * n1},  {(void**)&f->signs, 8 * n1},
        {(void**)&f->info, 64},       {(void**)&f->flags, 4 * nf},  {(void**)&f->prog, 4 * nf},
        {(void**)&f->ticket, 64},     {(void**)&f->hang, 64},       {(void**)&f->gmw, 64},
        {(void**)&f->cnt, 4 * (size_t)(H.dag_counters + 1)},        {(void**)&f->eacnt, 8 * nf},
        {(void**)&f->perm_dev, 8 * n1}};
    v.insert(v.end(), std::begin(rest), std::end(rest));
    return v;
}

static size_t arena_round(size_t b) { return (std::max<size_t>(b, 16) + 255) & ~(size_t)255; }

// Bytes of an arena factor of plan H (FP64).
static size_t arena_factor_bytes(const HostPlan& H) {
    mf_factor probe;
    size_t b = 0;
    for (auto& x : factor_buffers(&probe, H, 8, false)) b += arena_round(x.second);
    return b;
}

// Creates a factor object; arena != NULL: its buffers are carved from the
// arena (from arena->used on; nothing is freed with the object).
static int factor_create(mf_plan* pl, int precision, Arena* arena, mf_factor** out) {
    if (!pl || !out || (precision != MF_FP64 && precision != MF_FP32))
        return fail(MF_EARG, "bad argument");
    *out = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(MF_EHIP, "hipGetDevice failed");
    mf_factor* f = new mf_factor();
    f->plan = pl;
    f->precision = precision;
    f->device = dev;
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;
        f->stall = (unsigned long long)khz * 1000ull * 10ull;
    }
    const HostPlan& H = pl->H;
    size_t es = precision == MF_FP32 ? 4 : 8;
    auto alloc = [&](void** p, size_t bytes) -> bool {
        if (arena) {
            const size_t b = arena_round(bytes);
            if (arena->used + b > arena->size) return false;
            *p = arena->base + arena->used;
            arena->used += b;
            return true;
        }
        if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
        f->allocs.push_back(*p);
        f->alloc_bytes += std::max<size_t>(bytes, 16);
        return true;
    };
    bool ok = true;
    for (auto& x : factor_buffers(f, H, es, arena == nullptr)) ok = ok && alloc(x.first, x.second);
    if (!ok) {
        delete f;
        return fail(MF_ENOMEM, "hipMalloc failed (factor)");
    }
    if (hipStreamCreateWithFlags(&f->own, hipStreamNonBlocking) != hipSuccess) {
        delete f;
        return fail(MF_EHIP, "hipStreamCreate failed");
    }
    f->stream = f->own;
    launch_fill<int32_t>(f->hang, 16, 0, f->stream);
    if (H.n > 0) {
        if (precision == MF_FP32) launch_fill<float>((float*)f->d, H.n, 1.0f, f->stream);
        else launch_fill<double>((double*)f->d, H.n, 1.0, f->stream);
        launch_fill<double>(f->dref, H.n, 1.0, f->stream);
    }
    if (hipStreamSynchronize(f->stream) != hipSuccess) {
        delete f;
        return fail(MF_EHIP, "initialization failed");
    }
    *out = f;
    return MF_OK;
}

extern "C" {

int mf_factor_create(mf_plan* pl, int precision, mf_factor** out) {
    return factor_create(pl, precision, nullptr, out);
}

int mf_factor_destroy(mf_factor* f) {
    if (f) (void)hipStreamSynchronize(f->stream);
    delete f;
    return MF_OK;
}

int mf_factor_set_stream(mf_factor* f, void* stream) {
    if (!f) return fail(MF_EARG, "null factor");
    f->stream = stream ? (hipStream_t)stream : f->own;
    return MF_OK;
}

int mf_factor_set_solve_fp64(mf_factor* f, int on) {
    if (!f) return fail(MF_EARG, "null factor");
    if (f->precision != MF_FP32) return on ? fail(MF_EARG, "FP64 solves are for FP32 factor objects") : MF_OK;
    const HostPlan& H = f->plan->H;
    if (on && !f->sh_panel) {
        auto grab = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 16)) == hipSuccess; };
        if (!grab(&f->sh_panel, 8 * (size_t)H.panel_elems) || !grab(&f->sh_W, 8 * (size_t)H.w_elems) ||
            !grab(&f->sh_d, 8 * (size_t)(H.n + 1)))
            return fail(MF_ENOMEM, "hipMalloc failed (FP64 solve copy)");
        f->alloc_bytes += 8 * (size_t)(H.panel_elems + H.w_elems + H.n + 1);
    }
    f->solve64 = on != 0;
    f->sh_valid = false;
    if (f->solve64 && f->factored) enqueue_widen(f);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

int mf_factor_sync(mf_factor* f) {
    if (!f) return fail(MF_EARG, "null factor");
    HIPCHK(hipStreamSynchronize(f->stream));
    return MF_OK;
}

int mf_factor_set_values(mf_factor* f, const double* Dval, const double* Lval, int device) {
    if (!f || (!Dval && f->plan->H.ndz) || (!Lval && f->plan->H.nlz))
        return fail(MF_EARG, "null argument");
    const HostPlan& H = f->plan->H;
    hipMemcpyKind k = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (H.ndz) HIPCHK(hipMemcpyAsync(f->Dref, Dval, 8 * H.ndz, k, f->stream));
    if (H.nlz) HIPCHK(hipMemcpyAsync(f->Lref, Lval, 8 * H.nlz, k, f->stream));
    timed(f, MF_KC_ASSEMBLE, 8.0 * (double)(H.ndz + H.nlz), 1, [&] {
        if (f->precision == MF_FP32) {
            launch_fill<float>((float*)f->panel, H.panel_elems, 0.0f, f->stream);
            launch_assemble_ref<float>(f->plan->D, dev_factor<float>(f), f->Dref, f->Lref, f->uplo,
                                       f->adiag, f->stream);
        } else {
            launch_fill<double>((double*)f->panel, H.panel_elems, 0.0, f->stream);
            launch_assemble_ref<double>(f->plan->D, dev_factor<double>(f), f->Dref, f->Lref, f->uplo,
                                        f->adiag, f->stream);
        }
    });
    HIPCHK(hipGetLastError());
    if (!device) HIPCHK(hipStreamSynchronize(f->stream));
    f->factored = false;
    f->sel_valid = false;
    f->sh_valid = false;
    f->ref_stale = false;
    return MF_OK;
}

int mf_factor_load(mf_factor* f, const double* Dval, const double* Lval, const double* d, int diag,
                   int device) {
    if (!f || (diag != MF_CHOLESKY && diag != MF_LDLT)) return fail(MF_EARG, "bad argument");
    if (f->precision != MF_FP64) return fail(MF_EARG, "mf_factor_load needs an FP64 factor object");
    const HostPlan& H = f->plan->H;
    if (diag == MF_LDLT && !d && H.n) return fail(MF_EARG, "LDLt factor without d");
    int rc = mf_factor_set_values(f, Dval, Lval, device);
    if (rc) return rc;
    hipMemcpyKind k = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    f->diag = diag;
    if (H.n) {
        if (d) {
            HIPCHK(hipMemcpyAsync(f->dref, d, 8 * H.n, k, f->stream));
            if (diag == MF_LDLT) HIPCHK(hipMemcpyAsync(f->d, d, 8 * H.n, k, f->stream));
        } else {
            launch_fill<double>(f->dref, H.n, 1.0, f->stream);
        }
    }
    launch_fill<unsigned long long>(f->info, 1, ~0ull, f->stream);
    // W = inverses of the 64 x 64 diagonal blocks, as the factorization leaves them
    launch_diag<double>(f->plan->D, dev_factor<double>(f), 0, (int)H.diag.size(), true, f->stream);
    HIPCHK(hipGetLastError());
    if (!device) HIPCHK(hipStreamSynchronize(f->stream));
    f->factored = true;
    f->sel_valid = false;
    f->ref_stale = false;
    return MF_OK;
}

int mf_factor_set_pattern(mf_factor* f, int64_t nnz, const int64_t* flat) {
    if (!f || nnz < 0 || (nnz > 0 && !flat)) return fail(MF_EARG, "bad argument");
    const HostPlan& H = f->plan->H;
    std::vector<int64_t> tp(nnz, -1), td(nnz, -1);
    for (int64_t k = 0; k < nnz; ++k) {
        int64_t p = flat[k];
        if (p == 0) continue;
        if (p < 0 || p > H.ndz + H.nlz) return fail(MF_EARG, "flat index out of range");
        int64_t q = p - 1;
        int64_t j, e, r, c;
        const OutFront* o;
        if (q < H.ndz) {
            j = std::upper_bound(H.Dptr.begin(), H.Dptr.end(), q) - H.Dptr.begin() - 1;
            o = &H.out[j];
            e = q - H.Dptr[j];
            r = e % o->nn;
            c = e / o->nn;
            if (f->uplo == MF_UPLO_U) std::swap(r, c);   // :U stores the transposed block
            if (r < c) return fail(MF_EARG, "flat index outside the stored triangle");
            if (r == c) td[k] = o->r0 + c;
            const Front& F = H.fronts[o->J];
            int64_t cl = o->r0 - F.c0, nj = F.nn + F.na;
            tp[k] = F.poff + (cl + r) + (cl + c) * nj;
        } else {
            q -= H.ndz;
            j = std::upper_bound(H.Lptr.begin(), H.Lptr.end(), q) - H.Lptr.begin() - 1;
            // skip fronts with empty L blocks (Lptr repeats)
            while (j > 0 && H.Lptr[j] == H.Lptr[j + 1]) --j;
            o = &H.out[j];
            e = q - H.Lptr[j];
            if (f->uplo == MF_UPLO_U) {   // nn x na block above the diagonal block
                c = e % o->nn;
                r = e / o->nn;
            } else {
                r = e % o->na;
                c = e / o->na;
            }
            const Front& F = H.fronts[o->J];
            int64_t cl = o->r0 - F.c0, nj = F.nn + F.na;
            tp[k] = F.poff + H.septpos[o->sepoff + r] + (cl + c) * nj;
        }
    }
    if (f->tgt_panel) (void)hipFree(f->tgt_panel);
    if (f->tgt_diag) (void)hipFree(f->tgt_diag);
    f->tgt_panel = f->tgt_diag = nullptr;
    f->nnz = -1;
    HIPCHK(hipMalloc(&f->tgt_panel, 8 * std::max<int64_t>(nnz, 2)));
    HIPCHK(hipMalloc(&f->tgt_diag, 8 * std::max<int64_t>(nnz, 2)));
    if (nnz) {
        HIPCHK(hipMemcpy(f->tgt_panel, tp.data(), 8 * nnz, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(f->tgt_diag, td.data(), 8 * nnz, hipMemcpyHostToDevice));
    }
    f->nnz = nnz;
    return MF_OK;
}

int mf_factor_refresh(mf_factor* f, const double* nzval, int device) {
    if (!f || f->nnz < 0) return fail(MF_ESTATE, "no pattern registered");
    if (f->nnz > 0 && !nzval) return fail(MF_EARG, "null values");
    const HostPlan& H = f->plan->H;
    const double* src = nzval;
    if (!device && f->nnz > 0) {
        if (ensure((void**)&f->stage, &f->stage_bytes, 8 * (size_t)f->nnz)) return MF_ENOMEM;
        HIPCHK(hipMemcpyAsync(f->stage, nzval, 8 * f->nnz, hipMemcpyHostToDevice, f->stream));
        src = f->stage;
    }
    // fill!(F, 0) + setflatindex! into the panels only: the reference-layout
    // arrays are written from the panels when a caller asks for them
    timed(f, MF_KC_ASSEMBLE, (double)elt(f) * (double)H.panel_elems + 24.0 * f->nnz, 1, [&] {
        if (f->precision == MF_FP32) {
            launch_fill<float>((float*)f->panel, H.panel_elems, 0.0f, f->stream);
            launch_refresh<float>(dev_factor<float>(f), f->tgt_panel, f->tgt_diag, src, f->nnz,
                                  f->adiag, f->stream);
        } else {
            launch_fill<double>((double*)f->panel, H.panel_elems, 0.0, f->stream);
            launch_refresh<double>(dev_factor<double>(f), f->tgt_panel, f->tgt_diag, src, f->nnz,
                                   f->adiag, f->stream);
        }
    });
    if (f->A_val && f->A_nnz == f->nnz && f->nnz > 0)
        HIPCHK(hipMemcpyAsync(f->A_val, src, 8 * f->nnz, hipMemcpyDeviceToDevice, f->stream));
    HIPCHK(hipGetLastError());
    if (!device) HIPCHK(hipStreamSynchronize(f->stream));
    f->factored = false;
    f->sel_valid = false;
    f->sh_valid = false;
    f->ref_stale = true;
    return MF_OK;
}

int mf_factorize_async(mf_factor* f, int diag, const double* signs_dev, int reg_kind,
                       double delta, double eps) {
    if (!f || (diag != MF_CHOLESKY && diag != MF_LDLT) ||
        (reg_kind != MF_REG_NONE && reg_kind != MF_REG_DYNAMIC && reg_kind != MF_REG_GMW81))
        return fail(MF_EARG, "bad argument");
    // checksigns (no_regularization.jl:17-25): GMW81 needs signs +-1, which
    // LDLt's default (Zeros) is not
    if (reg_kind == MF_REG_GMW81 && diag == MF_LDLT && !signs_dev)
        return fail(MF_EARG, "GMW81 with LDLt needs signs");
    const HostPlan& H = f->plan->H;
    if (reg_kind == MF_REG_GMW81 && H.dag_gmw.size() != H.dag.size())
        return fail(MF_EARG, "this plan has no GMW81 order");
    f->diag = diag;
    f->reg_kind = reg_kind;
    f->delta = delta;
    f->eps = eps;
    if (reg_kind == MF_REG_GMW81 && delta < 0)  // gmw81_delta: cbrt(eps(T))
        f->delta = f->precision == MF_FP32 ? (double)std::cbrt(1.1920928955078125e-07f)
                                           : std::cbrt(2.220446049250313e-16);
    double* keep = f->signs;
    if (signs_dev) {
        if (signs_dev != f->signs)
            HIPCHK(hipMemcpyAsync(f->signs, signs_dev, 8 * H.n, hipMemcpyDeviceToDevice, f->stream));
    } else {
        f->signs = nullptr;  // default signs: Ones (Cholesky) / Zeros (LDLt)
    }
    if (f->precision == MF_FP32) enqueue_factor<float>(f);
    else enqueue_factor<double>(f);
    f->signs = keep;
    HIPCHK(hipGetLastError());
    f->factored = true;
    f->sel_valid = false;
    f->ref_stale = true;
    return MF_OK;
}

int mf_factor_info(mf_factor* f, int64_t* info) {
    if (!f) return fail(MF_EARG, "null factor");
    return read_info(f, info);
}

int mf_factorize(mf_factor* f, int diag, const double* signs, int reg_kind, double delta,
                 double eps, int64_t* info) {
    if (!f) return fail(MF_EARG, "null factor");
    int rc;
    if (signs) {
        HIPCHK(hipMemcpyAsync(f->signs, signs, 8 * f->plan->H.n, hipMemcpyHostToDevice, f->stream));
        rc = mf_factorize_async(f, diag, f->signs, reg_kind, delta, eps);
    } else {
        rc = mf_factorize_async(f, diag, nullptr, reg_kind, delta, eps);
    }
    if (rc) return rc;
    return read_info(f, info);
}

int mf_factor_get(mf_factor* f, double* Dval, double* Lval, double* d, int device) {
    if (!f) return fail(MF_EARG, "null factor");
    const HostPlan& H = f->plan->H;
    sync_ref(f);
    HIPCHK(hipGetLastError());
    hipMemcpyKind k = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (Dval && H.ndz) HIPCHK(hipMemcpyAsync(Dval, f->Dref, 8 * H.ndz, k, f->stream));
    if (Lval && H.nlz) HIPCHK(hipMemcpyAsync(Lval, f->Lref, 8 * H.nlz, k, f->stream));
    if (d && H.n) HIPCHK(hipMemcpyAsync(d, f->dref, 8 * H.n, k, f->stream));
    HIPCHK(hipStreamSynchronize(f->stream));
    return MF_OK;
}

int mf_solve_natural(mf_factor* f, double* C, int64_t nrhs, int64_t ldc, int device) {
    if (!f || nrhs < 0 || (nrhs > 0 && !C)) return fail(MF_EARG, "bad argument");
    const HostPlan& H = f->plan->H;
    if (H.tune.factor_only) return fail(MF_EARG, "this plan has no solve tasks");
    if (ldc < H.n) return fail(MF_EARG, "ldc < n");
    if (nrhs == 0 || H.n == 0) return MF_OK;
    if (solves64(f) && !f->sh_valid) return fail(MF_ESTATE, "FP64 solves of an FP32 factor: factorize after enabling them");
    size_t es = solve_elt(f);
    if (ensure(&f->Mv, &f->Mv_bytes, es * (size_t)H.mv_elems1 * nrhs)) return MF_ENOMEM;
    if ((f->precision == MF_FP64 || solves64(f)) && device) {
        enqueue_solve<double>(f, C, ldc, (int)nrhs);
    } else {
        if (ensure(&f->Bws, &f->Bws_bytes, es * (size_t)H.n * nrhs)) return MF_ENOMEM;
        // identity permutation copies through the workspace
        std::vector<int64_t> id;
        const double* src = C;
        double* stage = nullptr;
        if (!device) {
            if (ensure((void**)&f->stage, &f->stage_bytes, 8 * (size_t)H.n * nrhs)) return MF_ENOMEM;
            for (int64_t c = 0; c < nrhs; ++c)
                HIPCHK(hipMemcpyAsync(f->stage + c * H.n, C + c * ldc, 8 * H.n, hipMemcpyHostToDevice,
                                      f->stream));
            src = f->stage;
            stage = f->stage;
        }
        int64_t lds = device ? ldc : H.n;
        if (f->precision == MF_FP32 && !solves64(f)) {
            // permutation by the identity = conversion
            id.resize(H.n);
            for (int64_t i = 0; i < H.n; ++i) id[i] = i + 1;
            f->perm_host = id;
            HIPCHK(hipMemcpyAsync(f->perm_dev, f->perm_host.data(), 8 * H.n, hipMemcpyHostToDevice,
                                  f->stream));
            launch_permute<float>(src, lds, (float*)f->Bws, H.n, f->perm_dev, (int)H.n, (int)nrhs,
                                  false, f->stream);
            enqueue_solve<float>(f, (float*)f->Bws, H.n, (int)nrhs);
            double* dst = device ? C : stage;
            launch_permute_back(f->Bws, H.n, true, dst, device ? ldc : H.n, f->perm_dev, (int)H.n,
                                (int)nrhs, f->stream);
        } else {
            HIPCHK(hipMemcpyAsync(f->Bws, src, 8 * H.n * nrhs, hipMemcpyDeviceToDevice, f->stream));
            enqueue_solve<double>(f, (double*)f->Bws, H.n, (int)nrhs);
            HIPCHK(hipMemcpyAsync(stage, f->Bws, 8 * H.n * nrhs, hipMemcpyDeviceToDevice, f->stream));
        }
        if (!device) {
            for (int64_t c = 0; c < nrhs; ++c)
                HIPCHK(hipMemcpyAsync(C + c * ldc, f->stage + c * H.n, 8 * H.n, hipMemcpyDeviceToHost,
                                      f->stream));
        }
    }
    HIPCHK(hipGetLastError());
    if (!device) HIPCHK(hipStreamSynchronize(f->stream));
    return MF_OK;
}

int mf_solve(mf_factor* f, const int64_t* perm, double* B, int64_t nrhs, int64_t ldb, int device) {
    if (!f || !perm || nrhs < 0 || (nrhs > 0 && !B)) return fail(MF_EARG, "bad argument");
    const HostPlan& H = f->plan->H;
    if (H.tune.factor_only) return fail(MF_EARG, "this plan has no solve tasks");
    if (ldb < H.n) return fail(MF_EARG, "ldb < n");
    if (nrhs == 0 || H.n == 0) return MF_OK;
    if (solves64(f) && !f->sh_valid) return fail(MF_ESTATE, "FP64 solves of an FP32 factor: factorize after enabling them");
    size_t es = solve_elt(f);
    if (ensure(&f->Mv, &f->Mv_bytes, es * (size_t)H.mv_elems1 * nrhs)) return MF_ENOMEM;
    if (ensure(&f->Bws, &f->Bws_bytes, es * (size_t)H.n * nrhs)) return MF_ENOMEM;
    if (f->perm_host.size() != (size_t)H.n ||
        std::memcmp(f->perm_host.data(), perm, 8 * H.n) != 0) {
        f->perm_host.assign(perm, perm + H.n);
        HIPCHK(hipMemcpyAsync(f->perm_dev, f->perm_host.data(), 8 * H.n, hipMemcpyHostToDevice,
                              f->stream));
    }
    double* src = B;
    int64_t lds = ldb;
    if (!device) {
        if (ensure((void**)&f->stage, &f->stage_bytes, 8 * (size_t)H.n * nrhs)) return MF_ENOMEM;
        for (int64_t c = 0; c < nrhs; ++c)
            HIPCHK(hipMemcpyAsync(f->stage + c * H.n, B + c * ldb, 8 * H.n, hipMemcpyHostToDevice,
                                  f->stream));
        src = f->stage;
        lds = H.n;
    }
    if (f->precision == MF_FP32 && !solves64(f)) {
        launch_permute<float>(src, lds, (float*)f->Bws, H.n, f->perm_dev, (int)H.n, (int)nrhs, false,
                              f->stream);
        enqueue_solve<float>(f, (float*)f->Bws, H.n, (int)nrhs);
        launch_permute_back(f->Bws, H.n, true, src, lds, f->perm_dev, (int)H.n, (int)nrhs, f->stream);
    } else {
        launch_permute<double>(src, lds, (double*)f->Bws, H.n, f->perm_dev, (int)H.n, (int)nrhs,
                               false, f->stream);
        enqueue_solve<double>(f, (double*)f->Bws, H.n, (int)nrhs);
        launch_permute_back(f->Bws, H.n, false, src, lds, f->perm_dev, (int)H.n, (int)nrhs,
                            f->stream);
    }
    if (!device) {
        for (int64_t c = 0; c < nrhs; ++c)
            HIPCHK(hipMemcpyAsync(B + c * ldb, f->stage + c * H.n, 8 * H.n, hipMemcpyDeviceToHost,
                                  f->stream));
    }
    HIPCHK(hipGetLastError());
    if (!device) HIPCHK(hipStreamSynchronize(f->stream));
    return MF_OK;
}

int mf_solve_noalloc(mf_factor* f, const int64_t* perm, double* B, int64_t nrhs, int64_t ldb) {
    if (!f || !perm || nrhs < 0 || (nrhs > 0 && !B)) return fail(MF_EARG, "bad argument");
    const HostPlan& H = f->plan->H;
    if (H.tune.factor_only) return fail(MF_EARG, "this plan has no solve tasks");
    if (ldb < H.n) return fail(MF_EARG, "ldb < n");
    if (nrhs == 0 || H.n == 0) return MF_OK;
    if (solves64(f) && !f->sh_valid) return fail(MF_ESTATE, "FP64 solves of an FP32 factor: factorize after enabling them");
    const size_t es = solve_elt(f);
    const size_t mv_bytes = (es * (size_t)H.mv_elems1 * (size_t)nrhs + 255) & ~(size_t)255;
    const size_t need = mv_bytes + es * (size_t)H.n * (size_t)nrhs;
    if (need > MF_NOALLOC_SCRATCH_BYTES)
        return fail(MF_EARG, "mf_solve_noalloc: " + std::to_string(need) + " bytes of scratch needed, limit " +
                                 std::to_string((unsigned long long)MF_NOALLOC_SCRATCH_BYTES));
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev != f->device) return fail(MF_EARG, "the factor's device is not the current device");
    void* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    void* mv = base;
    void* bws = (unsigned char*)base + mv_bytes;
    if (f->perm_host.size() != (size_t)H.n ||
        std::memcmp(f->perm_host.data(), perm, 8 * H.n) != 0) {
        f->perm_host.assign(perm, perm + H.n);
        HIPCHK(hipMemcpyAsync(f->perm_dev, f->perm_host.data(), 8 * H.n, hipMemcpyHostToDevice,
                              f->stream));
    }
    if (f->precision == MF_FP32 && !solves64(f)) {
        launch_permute<float>(B, ldb, (float*)bws, H.n, f->perm_dev, (int)H.n, (int)nrhs, false,
                              f->stream);
        enqueue_solve<float>(f, (float*)bws, H.n, (int)nrhs, mv);
        launch_permute_back(bws, H.n, true, B, ldb, f->perm_dev, (int)H.n, (int)nrhs, f->stream);
    } else {
        launch_permute<double>(B, ldb, (double*)bws, H.n, f->perm_dev, (int)H.n, (int)nrhs, false,
                               f->stream);
        enqueue_solve<double>(f, (double*)bws, H.n, (int)nrhs, mv);
        launch_permute_back(bws, H.n, false, B, ldb, f->perm_dev, (int)H.n, (int)nrhs, f->stream);
    }
    HIPCHK(hipGetLastError());
    return MF_OK;
}

int mf_factor_set_matrix(mf_factor* f, int64_t n, const int64_t* colptr, const int64_t* rowval,
                         const double* nzval, int device) {
    if (!f || !colptr || n != f->plan->H.n) return fail(MF_EARG, "bad argument");
    const int64_t nnz = colptr[n] - 1;
    if (colptr[0] != 1 || nnz < 0 || (nnz > 0 && !rowval)) return fail(MF_EARG, "bad CSC pattern");
    for (int64_t j = 0; j < n; ++j)
        if (colptr[j + 1] < colptr[j]) return fail(MF_EARG, "bad CSC pattern");
    for (int64_t p = 0; p < nnz; ++p)
        if (rowval[p] < 1 || rowval[p] > n) return fail(MF_EARG, "row index out of range");
    if (f->A_colptr) (void)hipFree(f->A_colptr);
    if (f->A_rowval) (void)hipFree(f->A_rowval);
    if (f->A_val) (void)hipFree(f->A_val);
    f->A_colptr = nullptr;
    f->A_rowval = nullptr;
    f->A_val = nullptr;
    f->A_n = -1;
    if (hipMalloc((void**)&f->A_colptr, 8 * (n + 1)) != hipSuccess ||
        hipMalloc((void**)&f->A_rowval, 8 * std::max<int64_t>(nnz, 1)) != hipSuccess ||
        hipMalloc((void**)&f->A_val, 8 * std::max<int64_t>(nnz, 1)) != hipSuccess)
        return fail(MF_ENOMEM, "hipMalloc failed (matrix)");
    HIPCHK(hipMemcpyAsync(f->A_colptr, colptr, 8 * (n + 1), hipMemcpyHostToDevice, f->stream));
    if (nnz > 0)
        HIPCHK(hipMemcpyAsync(f->A_rowval, rowval, 8 * nnz, hipMemcpyHostToDevice, f->stream));
    if (nzval && nnz > 0)
        HIPCHK(hipMemcpyAsync(f->A_val, nzval, 8 * nnz,
                              device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, f->stream));
    HIPCHK(hipStreamSynchronize(f->stream));
    f->A_n = n;
    f->A_nnz = nnz;
    return MF_OK;
}

int mf_refine_segmented(mf_factor* f, const int64_t* perm, const double* B, int64_t ldb,
                        double* X, int64_t ldx, int64_t nrhs, int device, double tol, int maxit,
                        int64_t nseg, const int64_t* segptr, int* iters, double* relres) {
    if (!f || !perm || nrhs < 0 || (nrhs > 0 && (!B || !X)) || maxit < 0 || nseg < 1 || !segptr)
        return fail(MF_EARG, "bad argument");
    if (f->A_n < 0) return fail(MF_ESTATE, "no matrix registered (mf_factor_set_matrix)");
    if (!f->factored) return fail(MF_ESTATE, "factor not computed");
    const int64_t n = f->plan->H.n;
    if (ldb < n || ldx < n) return fail(MF_EARG, "leading dimension < n");
    if (segptr[0] != 0 || segptr[nseg] != n) return fail(MF_EARG, "segments must cover 0..n");
    for (int64_t g = 0; g < nseg; ++g)
        if (segptr[g + 1] < segptr[g]) return fail(MF_EARG, "segments must be increasing");
    if (nseg * nrhs > (int64_t)1 << 30 || nrhs > 65535) return fail(MF_EARG, "too many segments or columns");
    if (iters) *iters = 0;
    if (relres) *relres = 0;
    if (nrhs == 0 || n == 0) return MF_OK;
    const size_t cols = 8 * (size_t)n * nrhs, nn = (size_t)nseg * nrhs;
    if (ensure(&f->ir_ws, &f->ir_ws_bytes, 3 * cols + 16 * nn + 8 * (nseg + 1) + 64)) return MF_ENOMEM;
    double* Bd = (double*)f->ir_ws;
    double* Xd = Bd + (size_t)n * nrhs;
    double* Rd = Xd + (size_t)n * nrhs;
    double* nrm = Rd + (size_t)n * nrhs;          // |b|^2 then |r|^2, nseg x nrhs each
    int64_t* segd = (int64_t*)(nrm + 2 * nn);
    hipStream_t s = f->stream;
    HIPCHK(hipMemcpyAsync(segd, segptr, 8 * (nseg + 1), hipMemcpyHostToDevice, s));
    const hipMemcpyKind in = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    for (int64_t c = 0; c < nrhs; ++c)
        HIPCHK(hipMemcpyAsync(Bd + c * n, B + c * ldb, 8 * n, in, s));
    launch_colnorm2(segd, (int)nseg, Bd, n, (int)nrhs, nrm, s);
    launch_copy_cols(n, Bd, n, Xd, n, (int)nrhs, s);
    int rc = mf_solve(f, perm, Xd, nrhs, n, 1);
    if (rc) return rc;
    std::vector<double> h(2 * nn);
    double worst = 0;
    int it = 0;
    for (;;) {
        launch_residual(n, f->A_colptr, f->A_rowval, f->A_val, Xd, n, Bd, n, Rd, n, (int)nrhs, s);
        launch_colnorm2(segd, (int)nseg, Rd, n, (int)nrhs, nrm + nn, s);
        HIPCHK(hipMemcpyAsync(h.data(), nrm, 16 * nn, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        worst = 0;
        for (size_t k = 0; k < nn; ++k) {
            const double b2 = h[k], r2 = h[nn + k];
            const double rr = b2 > 0 ? std::sqrt(r2 / b2) : std::sqrt(r2);
            worst = std::max(worst, rr);
        }
        if (!(worst > tol) || it >= maxit) break;
        if ((rc = mf_solve(f, perm, Rd, nrhs, n, 1))) return rc;
        launch_axpy(n, Rd, n, Xd, n, (int)nrhs, s);
        ++it;
    }
    const hipMemcpyKind out = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    for (int64_t c = 0; c < nrhs; ++c)
        HIPCHK(hipMemcpyAsync(X + c * ldx, Xd + c * n, 8 * n, out, s));
    HIPCHK(hipStreamSynchronize(s));
    if (iters) *iters = it;
    if (relres) *relres = worst;
    return MF_OK;
}

int mf_refine_steps(mf_factor* f, const int64_t* perm, const double* B, int64_t ldb, double* X,
                    int64_t ldx, int64_t nrhs, int steps, int64_t nseg, const int64_t* segptr,
                    double* nrm2) {
    if (!f || !perm || nrhs < 0 || (nrhs > 0 && (!B || !X || !nrm2)) || steps < 0 || nseg < 1 ||
        !segptr)
        return fail(MF_EARG, "bad argument");
    if (f->A_n < 0) return fail(MF_ESTATE, "no matrix registered (mf_factor_set_matrix)");
    if (!f->factored) return fail(MF_ESTATE, "factor not computed");
    const int64_t n = f->plan->H.n;
    if (ldb < n || ldx < n) return fail(MF_EARG, "leading dimension < n");
    if (segptr[0] != 0 || segptr[nseg] != n) return fail(MF_EARG, "segments must cover 0..n");
    for (int64_t g = 0; g < nseg; ++g)
        if (segptr[g + 1] < segptr[g]) return fail(MF_EARG, "segments must be increasing");
    if (nseg * nrhs > (int64_t)1 << 30 || nrhs > 65535) return fail(MF_EARG, "too many segments or columns");
    if (nrhs == 0 || n == 0) return MF_OK;
    const size_t cols = 8 * (size_t)n * nrhs;
    if (ensure(&f->ir_ws, &f->ir_ws_bytes, 3 * cols + 16 * (size_t)nseg * nrhs + 8 * (nseg + 1) + 64))
        return MF_ENOMEM;
    double* Rd = (double*)f->ir_ws + 2 * (size_t)n * nrhs;
    int64_t* segd = (int64_t*)(Rd + (size_t)n * nrhs + 2 * (size_t)nseg * nrhs);
    hipStream_t s = f->stream;
    HIPCHK(hipMemcpyAsync(segd, segptr, 8 * (nseg + 1), hipMemcpyHostToDevice, s));
    launch_colnorm2(segd, (int)nseg, B, ldb, (int)nrhs, nrm2, s);
    launch_copy_cols(n, B, ldb, X, ldx, (int)nrhs, s);
    int rc = mf_solve(f, perm, X, nrhs, ldx, 1);
    for (int it = 0; !rc && it < steps; ++it) {
        launch_residual(n, f->A_colptr, f->A_rowval, f->A_val, X, ldx, B, ldb, Rd, n, (int)nrhs, s);
        if (!(rc = mf_solve(f, perm, Rd, nrhs, n, 1))) launch_axpy(n, Rd, n, X, ldx, (int)nrhs, s);
    }
    if (rc) return rc;
    launch_residual(n, f->A_colptr, f->A_rowval, f->A_val, X, ldx, B, ldb, Rd, n, (int)nrhs, s);
    launch_colnorm2(segd, (int)nseg, Rd, n, (int)nrhs, nrm2 + (size_t)nseg * nrhs, s);
    HIPCHK(hipGetLastError());
    return MF_OK;
}

int mf_refine(mf_factor* f, const int64_t* perm, const double* B, int64_t ldb, double* X,
              int64_t ldx, int64_t nrhs, int device, double tol, int maxit, int* iters,
              double* relres) {
    if (!f) return fail(MF_EARG, "null factor");
    const int64_t seg[2] = {0, f->plan->H.n};
    return mf_refine_segmented(f, perm, B, ldb, X, ldx, nrhs, device, tol, maxit, 1, seg, iters,
                               relres);
}

// The drop-in entry points keep one FP64 factor object per plan, created on
// first use, so repeated calls allocate nothing new.
static int dropin_factor(mf_plan* plan, int uplo, mf_factor** out) {
    if (!plan) return fail(MF_EARG, "null plan");
    if (uplo != MF_UPLO_L && uplo != MF_UPLO_U) return fail(MF_EARG, "bad uplo");
    if (!plan->dropin) {
        int rc = mf_factor_create(plan, MF_FP64, &plan->dropin);
        if (rc) return rc;
    }
    plan->dropin->uplo = uplo;
    plan->dropin->stream = plan->dropin->own;
    *out = plan->dropin;
    return MF_OK;
}

int mf_chol_impl(mf_plan* plan, double* Dval, double* Lval, double* d, const double* signs,
                 int diag, int uplo, int reg_kind, double delta, double eps, int64_t* info) {
    mf_factor* f = nullptr;
    int rc = dropin_factor(plan, uplo, &f);
    if (rc) return rc;
    if (!(rc = mf_factor_set_values(f, Dval, Lval, 0)) &&
        !(rc = mf_factorize(f, diag, signs, reg_kind, delta, eps, info)))
        rc = mf_factor_get(f, Dval, Lval, diag == MF_LDLT ? d : nullptr, 0);
    return rc;
}

int mf_div_impl(mf_plan* plan, const double* Dval, const double* Lval, const double* d,
                double* C, int64_t nrhs, int64_t ldc, int diag, int uplo) {
    if (diag != MF_CHOLESKY && diag != MF_LDLT) return fail(MF_EARG, "bad diag");
    if (diag == MF_LDLT && !d && plan && plan->H.n) return fail(MF_EARG, "LDLt solve without d");
    mf_factor* f = nullptr;
    int rc = dropin_factor(plan, uplo, &f);
    if (rc) return rc;
    f->diag = diag;
    if ((rc = mf_factor_set_values(f, Dval, Lval, 0))) return rc;
    if (diag == MF_LDLT && plan->H.n)
        HIPCHK(hipMemcpyAsync(f->d, d, 8 * plan->H.n, hipMemcpyHostToDevice, f->stream));
    // the big-front solves use the inverted 64x64 diagonal blocks
    launch_diag<double>(plan->D, dev_factor<double>(f), 0, (int)plan->H.diag.size(), true, f->stream);
    f->factored = true;
    f->sel_valid = false;
    f->ref_stale = false;
    return mf_solve_natural(f, C, nrhs, ldc, 0);
}

int mf_factor_bytes(const mf_factor* f, int64_t* bytes) {
    if (!f || !bytes) return fail(MF_EARG, "null argument");
    size_t b = f->alloc_bytes + f->Mv_bytes + f->Bws_bytes + f->ir_ws_bytes + f->stage_bytes;
    if (f->A_n >= 0) b += 8 * (size_t)(f->A_n + 1) + 16 * (size_t)std::max<int64_t>(f->A_nnz, 1);
    if (f->nnz >= 0) b += 16 * (size_t)std::max<int64_t>(f->nnz, 2);
    *bytes = (int64_t)b;
    return MF_OK;
}

int mf_factor_set_uplo(mf_factor* f, int uplo) {
    if (!f || (uplo != MF_UPLO_L && uplo != MF_UPLO_U)) return fail(MF_EARG, "bad argument");
    if (uplo != f->uplo) {
        f->nnz = -1;           // flat indices registered for the other layout no longer apply
        f->ref_stale = true;   // the reference arrays are rewritten in the new layout on demand
        f->uplo = uplo;
    }
    return MF_OK;
}

int mf_timing(mf_factor* f, int enable) {
    if (!f) return fail(MF_EARG, "null factor");
    f->timing = enable != 0;
    return MF_OK;
}

int mf_dag_profile(mf_factor* f, int enable) {
    if (!f) return fail(MF_EARG, "null factor");
    if (enable && !f->prof) {
        void* p = nullptr;
        if (hipMalloc(&p, 8 * (MF_DT_KINDS + 1) * 3) != hipSuccess) return fail(MF_ENOMEM, "hipMalloc failed (profile)");
        f->allocs.push_back(p);
        f->prof_buf = (unsigned long long*)p;
        launch_fill<unsigned long long>(f->prof_buf, (MF_DT_KINDS + 1) * 3, 0ull, f->stream);
        const size_t nt = f->plan->H.dag.size();
        if (hipMalloc(&p, 16 * (nt + 1)) != hipSuccess) return fail(MF_ENOMEM, "hipMalloc failed (trace)");
        f->allocs.push_back(p);
        f->trace_buf = (unsigned long long*)p;
        launch_fill<unsigned long long>(f->trace_buf, 2 * (nt + 1), 0ull, f->stream);
        const size_t ns = f->plan->H.ftask.size() + f->plan->H.btask.size();
        if (hipMalloc(&p, 16 * (ns + 1)) != hipSuccess) return fail(MF_ENOMEM, "hipMalloc failed (trace)");
        f->allocs.push_back(p);
        f->strace_buf = (unsigned long long*)p;
        launch_fill<unsigned long long>(f->strace_buf, 2 * (ns + 1), 0ull, f->stream);
    }
    f->prof = enable ? f->prof_buf : nullptr;
    return MF_OK;
}

int mf_dag_trace_read(mf_factor* f, unsigned long long* out, int64_t n) {
    if (!f || !out) return fail(MF_EARG, "null argument");
    const int64_t nt = (int64_t)f->plan->H.dag.size();
    if (n < 2 * nt) return fail(MF_EARG, "trace buffer too small (2 x DAG tasks)");
    if (!f->trace_buf) return fail(MF_ESTATE, "profiling is off");
    HIPCHK(hipMemcpyAsync(out, f->trace_buf, 16 * nt, hipMemcpyDeviceToHost, f->stream));
    HIPCHK(hipStreamSynchronize(f->stream));
    return MF_OK;
}

int mf_solve_trace_read(mf_factor* f, unsigned long long* out, int64_t n) {
    if (!f || !out) return fail(MF_EARG, "null argument");
    const int64_t ns = (int64_t)(f->plan->H.ftask.size() + f->plan->H.btask.size());
    if (n < 2 * ns) return fail(MF_EARG, "trace buffer too small (2 x solve tasks)");
    if (!f->strace_buf) return fail(MF_ESTATE, "profiling is off");
    HIPCHK(hipMemcpyAsync(out, f->strace_buf, 16 * ns, hipMemcpyDeviceToHost, f->stream));
    HIPCHK(hipStreamSynchronize(f->stream));
    return MF_OK;
}

int mf_plan_solve_tasks(const mf_plan* pl, int32_t* tasks, int64_t n) {
    if (!pl || !tasks) return fail(MF_EARG, "null argument");
    const HostPlan& H = pl->H;
    const size_t ns = H.ftask.size() + H.btask.size();
    if (n < 5 * (int64_t)ns) return fail(MF_EARG, "buffer too small (5 x solve tasks)");
    size_t t = 0;
    for (const std::vector<PTask>* tl : {&H.ftask, &H.btask})
        for (const PTask& k : *tl) {
            const int32_t v[5] = {k.kind, k.front, k.a, k.b, k.c};
            std::copy(v, v + 5, tasks + 5 * t++);
        }
    return MF_OK;
}

int mf_plan_fronts(const mf_plan* pl, int64_t* out, int64_t n) {
    if (!pl || !out) return fail(MF_EARG, "null argument");
    const HostPlan& H = pl->H;
    if (n < 5 * (int64_t)H.fronts.size()) return fail(MF_EARG, "buffer too small (5 x fronts)");
    for (size_t J = 0; J < H.fronts.size(); ++J) {
        const Front& F = H.fronts[J];
        const int64_t v[5] = {F.c0, F.nn, F.na, F.parent, (int64_t)pl->in_small[J]};
        std::copy(v, v + 5, out + 5 * J);
    }
    return MF_OK;
}

int mf_plan_dag(const mf_plan* pl, int32_t* tasks, int64_t n) {
    if (!pl || !tasks) return fail(MF_EARG, "null argument");
    const HostPlan& H = pl->H;
    if (n < 6 * (int64_t)H.dag.size()) return fail(MF_EARG, "buffer too small (6 x DAG tasks)");
    for (size_t t = 0; t < H.dag.size(); ++t) {
        const DagTask& d = H.dag[t];
        const int32_t v[6] = {d.kind, d.front, d.a, d.b, d.c, d.d};
        std::copy(v, v + 6, tasks + 6 * t);
    }
    return MF_OK;
}

int mf_plan_gemm(const mf_plan* pl, int32_t* out, int64_t n) {
    if (!pl || !out) return fail(MF_EARG, "null argument");
    const HostPlan& H = pl->H;
    if (n < 4 * (int64_t)H.gemm.size()) return fail(MF_EARG, "buffer too small (4 x GEMM tasks)");
    for (size_t t = 0; t < H.gemm.size(); ++t) {
        const GemmTask& g = H.gemm[t];
        const int32_t v[4] = {g.m, g.n, g.k, g.lower | (g.wide << 1) | ((int32_t)g.mode << 2)};
        std::copy(v, v + 4, out + 4 * t);
    }
    return MF_OK;
}

int mf_dag_profile_read(mf_factor* f, double* out, int reset) {
    if (!f || !out) return fail(MF_EARG, "null argument");
    for (int i = 0; i < (MF_DT_KINDS + 1) * 3; ++i) out[i] = 0;
    if (!f->prof_buf) return MF_OK;
    unsigned long long v[(MF_DT_KINDS + 1) * 3];
    HIPCHK(hipMemcpyAsync(v, f->prof_buf, sizeof(v), hipMemcpyDeviceToHost, f->stream));
    HIPCHK(hipStreamSynchronize(f->stream));
    int dev = 0, khz = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    const double us = khz > 0 ? 1e3 / khz : 0.0;
    for (int k = 0; k <= MF_DT_KINDS; ++k) {
        out[3 * k] = v[3 * k] * us;
        out[3 * k + 1] = v[3 * k + 1] * us;
        out[3 * k + 2] = k < MF_DT_KINDS ? (double)v[3 * k + 2] : v[3 * k + 2] * us;
    }
    if (reset) launch_fill<unsigned long long>(f->prof_buf, (MF_DT_KINDS + 1) * 3, 0ull, f->stream);
    return MF_OK;
}

int mf_timing_read(mf_factor* f, double* times_ms, int64_t* counts, double* work, int reset) {
    if (!f) return fail(MF_EARG, "null factor");
    drain_timing(f);
    for (int k = 0; k < MF_KC_COUNT; ++k) {
        if (times_ms) times_ms[k] = f->t_ms[k];
        if (counts) counts[k] = f->t_cnt[k];
        if (work) work[k] = f->t_work[k];
        if (reset) {
            f->t_ms[k] = 0;
            f->t_cnt[k] = 0;
            f->t_work[k] = 0;
        }
    }
    return MF_OK;
}

}  // extern "C"

// ---- selected inversion ------------------------------------------------------

namespace {

int ensure_sel_plan(mf_plan* pl) {
    if (pl->sel) return MF_OK;
    SelPlan* S = new SelPlan();
    std::string err = build_selinv(pl->H, *S);
    if (!err.empty()) {
        delete S;
        return fail(MF_EARG, "selected-inversion schedule: " + err);
    }
    int rc = upload(S->task, &pl->d_sel, pl->allocs);
    if (!rc) rc = upload(S->xoff, &pl->d_selx, pl->allocs);
    if (!rc) rc = upload(S->moff, &pl->d_selm, pl->allocs);
    if (rc) {
        delete S;
        return rc;
    }
    pl->sel = S;
    return MF_OK;
}

template <typename T>
DevSel<T> dev_sel(mf_factor* f) {
    const mf_plan* pl = f->plan;
    const SelPlan& S = *pl->sel;
    DevSel<T> D{};
    T* base = (T*)f->sel_pool;
    D.base[SP_PANEL] = (T*)f->panel;
    D.base[SP_W] = (T*)f->W;
    D.base[SP_X] = base;
    D.base[SP_T] = base + S.x_elems;
    D.base[SP_LT] = base + 2 * S.x_elems;
    D.base[SP_M] = D.base[SP_LT] + S.lt_elems;
    D.base[SP_SIG] = D.base[SP_M] + S.m_elems;
    D.task = pl->d_sel;
    D.xoff = pl->d_selx;
    D.moff = pl->d_selm;
    D.cnt = f->sel_cnt;
    D.ticket = f->ticket;
    D.hang = f->hang;
    D.stall = f->stall;
    D.ntask = (int32_t)S.task.size();
    return D;
}

template <typename T>
void enqueue_selinv(mf_factor* f) {
    mf_plan* pl = f->plan;
    hipStream_t s = f->stream;
    DevSel<T> D = dev_sel<T>(f);
    launch_fill<int32_t>(f->sel_cnt, (int64_t)pl->H.fronts.size() + 1, 0, s);
    launch_fill<uint32_t>(f->ticket, 1, 0u, s);
    timed(f, MF_KC_SELINV, pl->sel->flops, 1, [&] { launch_selinv<T>(pl->D, D, s); });
}

// The selected inverse into the factor object's reference-layout arrays
// (Dref/Lref), through the copy-out kernel with the Sigma panels in place of
// the factor's panels; the factor's own reference arrays are produced again
// from its panels when next read.
template <typename T>
void enqueue_sel_copyout(mf_factor* f) {
    DevFactor<T> F = dev_factor<T>(f);
    F.panel = dev_sel<T>(f).base[SP_SIG];
    F.diag = MF_CHOLESKY;
    launch_copyout<T>(f->plan->D, F, f->Dref, f->Lref, nullptr, f->uplo, f->adiag, f->stream);
    f->ref_stale = true;
}

}  // namespace

int mf_selinv(mf_factor* f) {
    if (!f) return fail(MF_EARG, "null factor");
    if (!f->factored) return fail(MF_ESTATE, "factor not computed");
    if (f->diag != MF_CHOLESKY)
        return fail(MF_EARG, "selected inversion needs a Cholesky factor (selinv!, selinv.jl:38-41)");
    mf_plan* pl = f->plan;
    int rc = ensure_sel_plan(pl);
    if (rc) return rc;
    const SelPlan& S = *pl->sel;
    const size_t es = elt(f);
    const size_t need = es * (size_t)(2 * S.x_elems + S.lt_elems + S.m_elems + pl->H.panel_elems);
    if (ensure(&f->sel_pool, &f->sel_bytes, need)) return MF_ENOMEM;
    if (!f->sel_cnt && hipMalloc((void**)&f->sel_cnt, 4 * (pl->H.fronts.size() + 1)) != hipSuccess)
        return fail(MF_ENOMEM, "hipMalloc failed (selected inversion)");
    if (f->precision == MF_FP32) enqueue_selinv<float>(f);
    else enqueue_selinv<double>(f);
    HIPCHK(hipGetLastError());
    f->sel_valid = true;
    return MF_OK;
}

int mf_selinv_get(mf_factor* f, double* Dval, double* Lval, int device) {
    if (!f) return fail(MF_EARG, "null factor");
    if (!f->sel_valid) return fail(MF_ESTATE, "no selected inverse computed (mf_selinv)");
    const HostPlan& H = f->plan->H;
    if (f->precision == MF_FP32) enqueue_sel_copyout<float>(f);
    else enqueue_sel_copyout<double>(f);
    HIPCHK(hipGetLastError());
    hipMemcpyKind k = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (Dval && H.ndz) HIPCHK(hipMemcpyAsync(Dval, f->Dref, 8 * H.ndz, k, f->stream));
    if (Lval && H.nlz) HIPCHK(hipMemcpyAsync(Lval, f->Lref, 8 * H.nlz, k, f->stream));
    HIPCHK(hipStreamSynchronize(f->stream));
    return read_info(f, nullptr);
}

// Elements of K per chunk of the completion's forest (a chunk's arena holds
// at most about four times that many doubles: the panels, the inverted
// diagonal tiles and the solves' Y and U); MF_COMPLETE_CHUNK_ELEMS overrides.
static int64_t cpl_budget() {
    if (const char* e = std::getenv("MF_COMPLETE_CHUNK_ELEMS")) {
        const long long v = std::atoll(e);
        if (v > 0) return (int64_t)v;
    }
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = (size_t)1 << 33;
    return std::max<int64_t>(std::min<int64_t>((int64_t)(fr / 64), (int64_t)1 << 30), (int64_t)1 << 16);
}

// Builds the completion's forest of a plan (on first use).
static int ensure_cpl(mf_plan* pl) {
    if (pl->cpl) return MF_OK;
    const HostPlan& H = pl->H;
    const int64_t nfr = H.nfr;
    CplState* C = new CplState();
    // K's separator rows: na, padded to a multiple of 64 in a big front
    auto sep_of = [&](int64_t j) {
        const int64_t nn = H.resptr[j + 1] - H.resptr[j], na = H.sepptr[j + 1] - H.sepptr[j];
        return nn + na > SMALL_NJ ? (na + 63) / 64 * 64 : na;
    };
    auto nj_of = [&](int64_t j) { return (H.resptr[j + 1] - H.resptr[j]) + sep_of(j); };
    const int64_t budget = cpl_budget();
    int rc = MF_OK;
    for (int64_t t0 = 0; t0 < nfr && !rc;) {
        int64_t t1 = t0, elems = 0;
        while (t1 < nfr && (t1 == t0 || elems + nj_of(nfr - 1 - t1) * nj_of(nfr - 1 - t1) <= budget)) {
            elems += nj_of(nfr - 1 - t1) * nj_of(nfr - 1 - t1);
            ++t1;
        }
        const int64_t m = t1 - t0;
        CplChunk c;
        c.t0 = t0;
        c.t1 = t1;
        std::vector<int64_t> resptr(m + 1, 1), ones(m + 2, 1), Dptr(m + 1, 1), dummy(1, 1);
        std::vector<int32_t> fr(m);
        c.sep.resize(m);
        c.sepp.resize(m);
        for (int64_t u = 0; u < m; ++u) {
            const int64_t j = nfr - 1 - (t0 + u), nj = nj_of(j);
            fr[u] = (int32_t)j;
            c.sep[u] = H.sepptr[j + 1] - H.sepptr[j];
            c.sepp[u] = sep_of(j);
            resptr[u + 1] = resptr[u] + nj;
            Dptr[u + 1] = Dptr[u] + nj * nj;
        }
        PlanTuning tune;
        tune.factor_only = true;
        rc = plan_create(resptr[m] - 1, m, resptr.data(), ones.data(), dummy.data(), dummy.data(), ones.data(),
                         dummy.data(), Dptr.data(), ones.data(), tune, &c.syn);
        if (rc) break;
        const HostPlan& SH = c.syn->H;
        const int32_t NJ = (int32_t)SH.fronts.size();
        std::vector<int64_t> poff(m);
        std::vector<int32_t> pad(NJ, 0), nnr(NJ, 0);
        for (int64_t u = 0; u < m; ++u) {
            const int32_t J = SH.owner[u];
            const Front& F = SH.fronts[J];
            // a root of the forest is its own internal front: nn = nJ, na = 0
            if (F.nn != resptr[u + 1] - resptr[u] || F.na != 0 || F.c0 != resptr[u] - 1) {
                rc = fail(MF_EARG, "completion forest: a root is not one internal front");
                break;
            }
            poff[u] = F.poff;
            pad[J] = (int32_t)c.sepp[u];
            nnr[J] = (int32_t)(F.nn - c.sepp[u]);
        }
        if (rc) break;
        SelPlan SP;
        const std::string err = build_cpl_solve(SH, pad, nnr, SP);
        if (!err.empty()) {
            rc = fail(MF_EARG, err);
            break;
        }
        std::vector<int64_t> xoff(m);
        for (int64_t u = 0; u < m; ++u) xoff[u] = SP.xoff[SH.owner[u]];
        c.ntask = (int32_t)SP.task.size();
        c.nctr = std::max<int64_t>(SP.ncounters, NJ) + 1;
        c.y_elems = SP.x_elems;
        if ((rc = upload(fr, &c.d_fr, C->allocs)) || (rc = upload(poff, &c.d_poff, C->allocs)) ||
            (rc = upload(xoff, &c.d_xoff, C->allocs)) || (c.ntask && (rc = upload(SP.task, &c.d_task, C->allocs))))
            break;
        // the arena holds the largest chunk's factor object, Y, U, R and counters
        const size_t need = arena_factor_bytes(SH) + 3 * arena_round(8 * (size_t)c.y_elems) +
                            arena_round(4 * (size_t)c.nctr);
        C->arena.size = std::max(C->arena.size, need);
        c.resptr = std::move(resptr);
        C->chunks.push_back(std::move(c));
        t0 = t1;
    }
    if (!rc) {
        std::vector<int32_t> septgt(H.septgt.begin(), H.septgt.end());
        std::vector<double> eye(64 * 64, 0.0);
        for (int i = 0; i < 64; ++i) eye[i + 64 * i] = 1.0;
        if (!(rc = upload(septgt, &C->d_septgt, C->allocs))) rc = upload(eye, &C->d_eye, C->allocs);
    }
    auto dalloc = [&](double** p, size_t n) {
        if (hipMalloc((void**)p, 8 * std::max<size_t>(n, 1)) != hipSuccess) return false;
        C->allocs.push_back(*p);
        return true;
    };
    if (!rc && (!dalloc(&C->d_sig, (size_t)(H.ndz + H.nlz)) || !dalloc(&C->d_out, (size_t)(H.ndz + H.nlz)) ||
                !dalloc((double**)&C->arena.base, C->arena.size / 8 + 1)))
        rc = fail(MF_ENOMEM, "hipMalloc failed (completion)");
    if (rc) {
        delete C;
        return rc;
    }
    pl->cpl = C;
    return MF_OK;
}

// Dval/Lval blocks transposed per front, between the :U and :L layouts
// (either direction: to_upper = false reads :U into :L, true writes :L as :U).
static void cpl_transpose(const HostPlan& H, const double* Din, const double* Lin, double* Dout, double* Lout,
                          bool to_upper) {
    for (int64_t j = 0; j < H.nfr; ++j) {
        const int64_t nn = H.resptr[j + 1] - H.resptr[j], na = H.sepptr[j + 1] - H.sepptr[j];
        const double* Di = Din + H.Dptr[j];
        double* Do = Dout + H.Dptr[j];
        for (int64_t c = 0; c < nn; ++c)
            for (int64_t r = 0; r < nn; ++r) Do[r + c * nn] = Di[c + r * nn];
        const double* Li = Lin + H.Lptr[j];
        double* Lo = Lout + H.Lptr[j];
        for (int64_t c = 0; c < nn; ++c)
            for (int64_t r = 0; r < na; ++r) {
                if (to_upper) Lo[c + r * nn] = Li[r + c * na];  // nn x na <- na x nn
                else Lo[r + c * na] = Li[c + r * nn];           // na x nn <- nn x na
            }
    }
}

// One chunk of the completion: K into the chunk factor's panels, Cholesky
// in place, the backward solves Y of the big roots on k_selinv, the outputs
// of the chunk's fronts into d_out.  *inf: the forest's failing column
// (1-based in the chunk), 0 if none.
static int cpl_chunk(mf_plan* plan, CplChunk& c, hipStream_t caller, int64_t* inf) {
    CplState& C = *plan->cpl;
    const HostPlan& H = plan->H;
    mf_factor* sf = nullptr;
    C.arena.used = 0;
    int rc = factor_create(c.syn, MF_FP64, &C.arena, &sf);
    if (rc) return rc;
    const int m = (int)(c.t1 - c.t0);
    // Sigma was copied on the caller's stream
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ev, caller));
    HIPCHK(hipStreamWaitEvent(sf->stream, ev, 0));
    HIPCHK(hipEventDestroy(ev));
    launch_cpl_gather(plan->D, m, c.d_fr, C.d_sig, C.d_sig + H.ndz, C.d_septgt, c.d_poff, (double*)sf->panel,
                      sf->stream);
    HIPCHK(hipGetLastError());
    sf->factored = false;
    sf->sel_valid = false;
    sf->sh_valid = false;
    sf->ref_stale = false;
    rc = mf_factorize(sf, MF_CHOLESKY, nullptr, MF_REG_NONE, 0.0, 0.0, inf);
    double* Y = (double*)(C.arena.base + C.arena.used);
    if (!rc && !*inf && c.ntask) {
        double* U = (double*)((char*)Y + arena_round(8 * (size_t)c.y_elems));
        double* R = (double*)((char*)U + arena_round(8 * (size_t)c.y_elems));
        int32_t* cnt = (int32_t*)((char*)R + arena_round(8 * (size_t)c.y_elems));
        DevSel<double> D{};
        D.base[SP_PANEL] = (double*)sf->panel;
        D.base[SP_W] = (double*)sf->W;
        D.base[SP_X] = Y;
        D.base[SP_T] = U;
        D.base[SP_LT] = C.d_eye;
        D.base[SP_M] = R;
        D.task = c.d_task;
        D.xoff = c.d_xoff;
        D.cnt = cnt;
        D.ticket = sf->ticket;
        D.hang = sf->hang;
        D.stall = sf->stall;
        D.ntask = c.ntask;
        launch_fill<int32_t>(cnt, c.nctr, 0, sf->stream);
        launch_fill<uint32_t>(sf->ticket, 1, 0u, sf->stream);
        launch_selinv<double>(c.syn->D, D, sf->stream);
        if (hipGetLastError() != hipSuccess) rc = fail(MF_EHIP, "completion: launch failed");
        if (!rc) rc = read_info(sf, nullptr);  // the solves' watchdog
    }
    if (!rc && !*inf) {
        launch_cpl_extract(plan->D, m, c.d_fr, Y, c.ntask ? c.d_xoff : nullptr, (const double*)sf->panel, c.d_poff,
                           C.d_out, C.d_out + H.ndz, sf->stream);
        if (hipGetLastError() != hipSuccess) rc = fail(MF_EHIP, "completion: launch failed");
    }
    if (hipStreamSynchronize(sf->stream) != hipSuccess && !rc) rc = fail(MF_EHIP, "completion: stream failed");
    (void)mf_factor_destroy(sf);
    return rc;
}

int mf_complete_impl(mf_plan* plan, double* Dval, double* Lval, int uplo, int64_t* info) {
    if (!plan || (!Dval && plan->H.ndz) || (!Lval && plan->H.nlz)) return fail(MF_EARG, "null argument");
    if (uplo != MF_UPLO_L && uplo != MF_UPLO_U) return fail(MF_EARG, "bad uplo");
    const HostPlan& H = plan->H;
    int rc = ensure_cpl(plan);
    if (rc) return rc;
    CplState& C = *plan->cpl;
    const hipStream_t s = nullptr;  // the null stream: transfers of Sigma and the result
    std::vector<double> tD, tL;
    const double *Din = Dval, *Lin = Lval;
    if (uplo == MF_UPLO_U) {
        tD.resize(H.ndz);
        tL.resize(H.nlz);
        cpl_transpose(H, Dval, Lval, tD.data(), tL.data(), false);
        Din = tD.data();
        Lin = tL.data();
    }
    if (H.ndz) HIPCHK(hipMemcpyAsync(C.d_sig, Din, 8 * H.ndz, hipMemcpyHostToDevice, s));
    if (H.nlz) HIPCHK(hipMemcpyAsync(C.d_sig + H.ndz, Lin, 8 * H.nlz, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    for (CplChunk& c : C.chunks) {
        int64_t inf = 0;
        if ((rc = cpl_chunk(plan, c, s, &inf))) return rc;
        if (inf) {
            // the first failing column of the chunk (its roots in the
            // reference's order): the local column of chol(Sigma(sep, sep)),
            // or of the reordered Schur complement
            const int64_t col = inf - 1;
            const int64_t u = std::upper_bound(c.resptr.begin(), c.resptr.end(), col + 1) - c.resptr.begin() - 1;
            const int64_t k = col - (c.resptr[u] - 1);  // padding columns (na <= k < na') cannot fail
            if (info) *info = k < c.sep[u] ? k + 1 : k - c.sepp[u] + 1;
            return MF_OK;
        }
    }
    if (uplo == MF_UPLO_U) {
        if (H.ndz) HIPCHK(hipMemcpy(tD.data(), C.d_out, 8 * H.ndz, hipMemcpyDeviceToHost));
        if (H.nlz) HIPCHK(hipMemcpy(tL.data(), C.d_out + H.ndz, 8 * H.nlz, hipMemcpyDeviceToHost));
        cpl_transpose(H, tD.data(), tL.data(), Dval, Lval, true);
    } else {
        if (H.ndz) HIPCHK(hipMemcpy(Dval, C.d_out, 8 * H.ndz, hipMemcpyDeviceToHost));
        if (H.nlz) HIPCHK(hipMemcpy(Lval, C.d_out + H.ndz, 8 * H.nlz, hipMemcpyDeviceToHost));
    }
    if (info) *info = 0;
    return MF_OK;
}

int mf_selinv_impl(mf_plan* plan, double* Dval, double* Lval, int uplo, int64_t* info) {
    if ((!Dval && plan && plan->H.ndz) || (!Lval && plan && plan->H.nlz)) return fail(MF_EARG, "null argument");
    mf_factor* f = nullptr;
    int rc = dropin_factor(plan, uplo, &f);
    if (rc) return rc;
    if ((rc = mf_factor_load(f, Dval, Lval, nullptr, MF_CHOLESKY, 0))) return rc;
    if ((rc = mf_selinv(f))) return rc;
    if ((rc = mf_selinv_get(f, Dval, Lval, 0))) return rc;
    if (info) *info = 0;
    return MF_OK;
}
