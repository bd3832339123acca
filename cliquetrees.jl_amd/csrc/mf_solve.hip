// mf_solve.hip — gfx950 kernels of the supernodal triangular solves.
//
// Forward  C <- L^-1 C  (div_impl! forward, divide.jl:417-450; per front:
//   F <- sum_i R_i M_i (div_fwd_update!, divide.jl:1206-1248), C1 += F1,
//   C1 <- D11^-1 C1, M2 <- F2 - L21 C1, divide.jl:493-659).
// Backward C <- L^-T C (divide.jl:864-1038): C1 -= L21' M2, where M2 is the
//   gather of the already final rows sep(j) of C (div_bwd_update!,
//   divide.jl:1250-1293), then C1 <- D11^-T C1.
//
// Each direction is ONE persistent launch over the whole front tree.
// Workgroups take tasks by ticket in a topological order (so every wait
// targets a task that a running workgroup already holds):
//   PT_SMALL  a whole small front (nJ <= 64) with its panel in LDS;
//   PT_DIAG   a 64-row diagonal block of a big front: gather of the children's
//             update vectors, the left-looking sum over the earlier blocks
//             (waiting on the front's progress counter block by block), and
//             the product with the inverted diagonal block W;
//   PT_ROWS   64 update rows of a big front: M2 rows = gather - L21 X.
// Completion: per-front task counters (children complete / parent complete)
// and per-front block progress counters.  Thread 0 polls with relaxed loads
// and issues one agent-scope acquire; after the closing barrier (every wave
// first drains its stores to L2) thread 0 writes L2 back, waits for the
// write-back, and only then publishes the task with relaxed atomics.
// All waits are bounded.  Right-hand sides are processed 16 at a time.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mf_kernels.h"

namespace mf {

namespace {

constexpr int RB = 16;  // right-hand sides per pass
constexpr int LDS = SMALL_NJ + 1;
constexpr int LDP = NB + 1;

// Waits until this wave's global stores (and loads) have completed, i.e. reached
// L2.  Before the closing barrier of a task every wave drains, so that thread
// 0's agent-scope release (an L2 write-back) covers the whole workgroup's data.
__device__ __forceinline__ void drain_stores() { __builtin_amdgcn_s_waitcnt(0); }
// Agent-scope acquire (L1/L2 invalidate) and release (L2 write-back), each
// followed by an explicit wait: the compiler omits the wait after the
// write-back when its own counters are already zero (after drain_stores), and
// a flag stored before the write-back completes can be seen by a workgroup on
// another XCD ahead of the data.  The wait after the invalidate makes it
// complete before the barrier releases the workgroup's other waves.
__device__ __forceinline__ void acquire_agent() {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
}
__device__ __forceinline__ void release_agent() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
}

// Bounded wait until *p >= target; gives up (and records it) after ~1 s.
__device__ __forceinline__ void wait_geq(int32_t* p, int32_t target, int32_t* hang) {
    int spins = 0;
    // poll with relaxed loads (an acquire per poll would invalidate the caches
    // every other workgroup is using); the caller issues one acquire fence
    while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++spins > (1 << 24)) {
            atomicExch(hang, 1);
            break;
        }
        // once any wait has given up, every other wait returns at once
        if ((spins & 255) == 0 && __hip_atomic_load(hang, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            break;
        __builtin_amdgcn_s_sleep(2);
    }
}

struct Sync {
    int32_t* done;    // per front: finished tasks
    int32_t* prog;    // per front: finished diagonal blocks (in chain order)
    const int32_t* nt;
    uint32_t* ticket;
};

// Children's update vectors scattered into Y (rows [r0, r1) of front f,
// column c of the pass at Y[(row - r0) + c * NB]).
template <typename T>
__device__ void gather_children(const DevPlan& P, const DevSolve<T>& S, const Front& f, int r0,
                                int r1, int cb, int nb, T* Y) {
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    for (int e = tid; e < NB * RB; e += blockDim.x) Y[e] = (T)0;
    __syncthreads();
    for (int q = f.chd0; q < f.chd1; ++q) {
        const int K = P.chd[q];
        const Front kf = P.fronts[K];
        const int ma = kf.na;
        const int32_t* relK = P.rel + kf.reloff;
        int lo = 0, hi = ma;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (relK[mid] < r0) lo = mid + 1; else hi = mid;
        }
        const int a = lo;
        hi = ma;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (relK[mid] < r1) lo = mid + 1; else hi = mid;
        }
        const T* MvK = S.Mv + P.mvoff1[K] * S.nrhs;
        for (int i = a + r; i < lo; i += 64) {
            const int yr = relK[i] - r0;
            for (int c = g; c < nb; c += 4) Y[yr + c * NB] += MvK[i + (int64_t)(cb + c) * ma];
        }
        __syncthreads();
    }
}

template <typename T>
__device__ void small_fwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                          const Front& f, int J, T* Ps, T* X) {
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    const int nn = f.nn, na = f.na, nj = nn + na;
    const bool unit = F.diag != 0;
    const T* Pj = F.panel + f.poff;
    if (r < nj)
        for (int c = g; c < nn; c += 4) Ps[r + c * LDS] = Pj[r + c * nj];
    T* MvJ = S.Mv + P.mvoff1[J] * S.nrhs;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        for (int c = g; c < RB; c += 4) X[r + c * LDS] = (T)0;
        __syncthreads();
        for (int q = f.chd0; q < f.chd1; ++q) {
            const int K = P.chd[q];
            const Front kf = P.fronts[K];
            const int ma = kf.na;
            if (r < ma) {
                const int rr = P.rel[kf.reloff + r];
                const T* MvK = S.Mv + P.mvoff1[K] * S.nrhs;
                for (int c = g; c < nb; c += 4) X[rr + c * LDS] += MvK[r + (int64_t)(cb + c) * ma];
            }
            __syncthreads();
        }
        if (r < nn)
            for (int c = g; c < nb; c += 4) X[r + c * LDS] += S.B[f.c0 + r + (int64_t)(cb + c) * S.ldb];
        __syncthreads();
        for (int k = 0; k < nn; ++k) {
            if (!unit && tid < nb) X[k + tid * LDS] /= Ps[k + k * LDS];
            __syncthreads();
            if (r > k && r < nj) {
                const T p = Ps[r + k * LDS];
                for (int c = g; c < nb; c += 4) X[r + c * LDS] -= p * X[k + c * LDS];
            }
            __syncthreads();
        }
        if (r < nj) {
            for (int c = g; c < nb; c += 4) {
                if (r < nn) S.B[f.c0 + r + (int64_t)(cb + c) * S.ldb] = X[r + c * LDS];
                else MvJ[(r - nn) + (int64_t)(cb + c) * na] = X[r + c * LDS];
            }
        }
        __syncthreads();
    }
}

template <typename T>
__device__ void small_bwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                          const Front& f, T* Ps, T* X) {
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    const int nn = f.nn, na = f.na, nj = nn + na;
    const bool unit = F.diag != 0;
    const T* Pj = F.panel + f.poff;
    if (r < nj)
        for (int c = g; c < nn; c += 4) Ps[r + c * LDS] = Pj[r + c * nj];
    const int32_t* rows = P.rows + f.rowoff;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        if (r < nj)
            for (int c = g; c < nb; c += 4) X[r + c * LDS] = S.B[rows[r] + (int64_t)(cb + c) * S.ldb];
        __syncthreads();
        if (r < nn) {
            for (int c = g; c < nb; c += 4) {
                T s = (T)0;
                for (int q = nn; q < nj; ++q) s += Ps[q + r * LDS] * X[q + c * LDS];
                X[r + c * LDS] -= s;
            }
        }
        (void)na;
        __syncthreads();
        for (int k = nn - 1; k >= 0; --k) {
            if (!unit && tid < nb) X[k + tid * LDS] /= Ps[k + k * LDS];
            __syncthreads();
            if (r < k) {
                const T p = Ps[k + r * LDS];
                for (int c = g; c < nb; c += 4) X[r + c * LDS] -= p * X[k + c * LDS];
            }
            __syncthreads();
        }
        if (r < nn)
            for (int c = g; c < nb; c += 4) S.B[f.c0 + r + (int64_t)(cb + c) * S.ldb] = X[r + c * LDS];
        __syncthreads();
    }
}

// Forward diagonal block [k0, k0+kb) of a big front (rows r = tid & 63,
// right-hand sides g, g+4, g+8, g+12 of the pass).
template <typename T>
__device__ void diag_fwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                         const Front& f, int J, int k0, int kb, const Sync& Y_, T* Ps, T* X,
                         T* Y) {
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    const int nj = f.nn + f.na;
    const T* Pj = F.panel + f.poff;
    const T* W = F.W + f.woff;
    const int kblk = k0 / NB;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        gather_children(P, S, f, k0, k0 + kb, cb, nb, Y);
        T acc[RB / 4];
#pragma unroll
        for (int i = 0; i < RB / 4; ++i) {
            const int c = g + 4 * i;
            acc[i] = (r < kb && c < nb) ? S.B[f.c0 + k0 + r + (int64_t)(cb + c) * S.ldb] + Y[r + c * NB]
                                        : (T)0;
        }
        for (int ib = 0; ib < kblk; ++ib) {
            if (tid == 0) {
                wait_geq(Y_.prog + J, ib + 1, F.hang);
                acquire_agent();
            }
            __syncthreads();
            for (int c = g; c < nb; c += 4) X[r + c * NB] = S.B[f.c0 + ib * NB + r + (int64_t)(cb + c) * S.ldb];
            for (int c = g; c < NB; c += 4) Ps[r + c * LDP] = (r < kb) ? Pj[(k0 + r) + (int64_t)(ib * NB + c) * nj] : (T)0;
            __syncthreads();
            for (int kk = 0; kk < NB; ++kk) {
                const T lv = Ps[r + kk * LDP];
#pragma unroll
                for (int i = 0; i < RB / 4; ++i) acc[i] -= lv * X[kk + (g + 4 * i) * NB];
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < RB / 4; ++i) Y[r + (g + 4 * i) * NB] = acc[i];
        for (int c = g; c < NB; c += 4) Ps[r + c * LDP] = (r < kb && c < kb) ? W[k0 + r + (int64_t)c * f.nn] : (T)0;
        __syncthreads();
        if (r < kb) {
#pragma unroll
            for (int i = 0; i < RB / 4; ++i) {
                const int c = g + 4 * i;
                if (c < nb) {
                    T s = (T)0;
                    for (int kk = 0; kk < kb; ++kk) s += Ps[r + kk * LDP] * Y[kk + c * NB];
                    S.B[f.c0 + k0 + r + (int64_t)(cb + c) * S.ldb] = s;
                }
            }
        }
        __syncthreads();
    }
}

// Forward update rows [r0, r1) (>= nn) of a big front: M2 = gather - L21 X.
template <typename T>
__device__ void rows_fwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                         const Front& f, int J, int r0, int r1, const Sync& Y_, T* Ps, T* X,
                         T* Y) {
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    const int nn = f.nn, na = f.na, nj = nn + na, m = r1 - r0;
    const T* Pj = F.panel + f.poff;
    const int p = (nn + NB - 1) / NB;
    T* MvJ = S.Mv + P.mvoff1[J] * S.nrhs;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        gather_children(P, S, f, r0, r1, cb, nb, Y);
        T acc[RB / 4];
#pragma unroll
        for (int i = 0; i < RB / 4; ++i) acc[i] = Y[r + (g + 4 * i) * NB];
        for (int ib = 0; ib < p; ++ib) {
            const int kbi = min(NB, nn - ib * NB);
            if (tid == 0) {
                wait_geq(Y_.prog + J, ib + 1, F.hang);
                acquire_agent();
            }
            __syncthreads();
            for (int c = g; c < nb; c += 4)
                X[r + c * NB] = (r < kbi) ? S.B[f.c0 + ib * NB + r + (int64_t)(cb + c) * S.ldb] : (T)0;
            for (int c = g; c < NB; c += 4)
                Ps[r + c * LDP] = (r < m && c < kbi) ? Pj[(r0 + r) + (int64_t)(ib * NB + c) * nj] : (T)0;
            __syncthreads();
            for (int kk = 0; kk < kbi; ++kk) {
                const T lv = Ps[r + kk * LDP];
#pragma unroll
                for (int i = 0; i < RB / 4; ++i) acc[i] -= lv * X[kk + (g + 4 * i) * NB];
            }
            __syncthreads();
        }
        if (r < m) {
#pragma unroll
            for (int i = 0; i < RB / 4; ++i) {
                const int c = g + 4 * i;
                if (c < nb) MvJ[(r0 + r - nn) + (int64_t)(cb + c) * na] = acc[i];
            }
        }
        __syncthreads();
    }
}

// Backward diagonal block [k0, k0+kb): columns r = tid & 63 of the block.
template <typename T>
__device__ void diag_bwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                         const Front& f, int J, int k0, int kb, const Sync& Y_, T* Ps, T* X,
                         T* Y) {
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    const int nn = f.nn, nj = nn + f.na;
    const T* Pj = F.panel + f.poff;
    const T* W = F.W + f.woff;
    const int32_t* rows = P.rows + f.rowoff;
    const int p = (nn + NB - 1) / NB, kblk = k0 / NB;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        T acc[RB / 4];
#pragma unroll
        for (int i = 0; i < RB / 4; ++i) {
            const int c = g + 4 * i;
            acc[i] = (r < kb && c < nb) ? S.B[f.c0 + k0 + r + (int64_t)(cb + c) * S.ldb] : (T)0;
        }
        // C1 -= L21' x_sep (x_sep final: ancestors are complete)
        for (int rt = nn; rt < nj; rt += NB) {
            const int nr = min(NB, nj - rt);
            for (int c = g; c < nb; c += 4)
                X[r + c * NB] = (r < nr) ? S.B[rows[rt + r] + (int64_t)(cb + c) * S.ldb] : (T)0;
            for (int c = g; c < NB; c += 4)
                Ps[r + c * LDP] = (r < nr && c < kb) ? Pj[(rt + r) + (int64_t)(k0 + c) * nj] : (T)0;
            __syncthreads();
            for (int q = 0; q < nr; ++q) {
                const T lv = Ps[q + r * LDP];
#pragma unroll
                for (int i = 0; i < RB / 4; ++i) acc[i] -= lv * X[q + (g + 4 * i) * NB];
            }
            __syncthreads();
        }
        // blocks below the diagonal block, finished in decreasing order
        for (int ib = p - 1; ib > kblk; --ib) {
            const int kbi = min(NB, nn - ib * NB);
            if (tid == 0) {
                wait_geq(Y_.prog + J, p - ib, F.hang);
                acquire_agent();
            }
            __syncthreads();
            for (int c = g; c < nb; c += 4)
                X[r + c * NB] = (r < kbi) ? S.B[f.c0 + ib * NB + r + (int64_t)(cb + c) * S.ldb] : (T)0;
            for (int c = g; c < NB; c += 4)
                Ps[r + c * LDP] = (r < kbi && c < kb) ? Pj[(ib * NB + r) + (int64_t)(k0 + c) * nj] : (T)0;
            __syncthreads();
            for (int q = 0; q < kbi; ++q) {
                const T lv = Ps[q + r * LDP];
#pragma unroll
                for (int i = 0; i < RB / 4; ++i) acc[i] -= lv * X[q + (g + 4 * i) * NB];
            }
            __syncthreads();
        }
        // X_k = W_k' acc
#pragma unroll
        for (int i = 0; i < RB / 4; ++i) Y[r + (g + 4 * i) * NB] = acc[i];
        for (int c = g; c < NB; c += 4) Ps[r + c * LDP] = (r < kb && c < kb) ? W[k0 + r + (int64_t)c * f.nn] : (T)0;
        __syncthreads();
        if (r < kb) {
#pragma unroll
            for (int i = 0; i < RB / 4; ++i) {
                const int c = g + 4 * i;
                if (c < nb) {
                    T s = (T)0;
                    for (int q = 0; q < kb; ++q) s += Ps[q + r * LDP] * Y[q + c * NB];
                    S.B[f.c0 + k0 + r + (int64_t)(cb + c) * S.ldb] = s;
                }
            }
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fsolve(DevPlan P, DevFactor<T> F, DevSolve<T> S,
                                                const PTask* tasks, int count, Sync Y_) {
    __shared__ T Ps[SMALL_NJ * LDS];
    __shared__ T X[LDS * RB];
    __shared__ T Y[NB * RB];
    __shared__ int s_t;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_t = (int)atomicAdd(Y_.ticket, 1u);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= count) break;
        const PTask tk = tasks[t];
        const int J = tk.front;
        const Front f = P.fronts[J];
        if (tid == 0) {
            for (int q = f.chd0; q < f.chd1; ++q) {
                const int K = P.chd[q];
                wait_geq(Y_.done + K, Y_.nt[K], F.hang);
            }
            acquire_agent();
        }
        __syncthreads();
        if (tk.kind == PT_SMALL) small_fwd(P, F, S, f, J, Ps, X);
        else if (tk.kind == PT_DIAG) diag_fwd(P, F, S, f, J, tk.a, tk.b, Y_, Ps, X, Y);
        else rows_fwd(P, F, S, f, J, tk.a, tk.b, Y_, Ps, X, Y);
        drain_stores();
        __syncthreads();
        if (tid == 0) {
            release_agent();
            if (tk.kind == PT_DIAG)
                __hip_atomic_store(Y_.prog + J, tk.a / NB + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(Y_.done + J, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // closes the thread-0 block before the back-edge: every thread of the
        // workgroup reaches the next ticket together
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bsolve(DevPlan P, DevFactor<T> F, DevSolve<T> S,
                                                const PTask* tasks, int count, Sync Y_) {
    __shared__ T Ps[SMALL_NJ * LDS];
    __shared__ T X[LDS * RB];
    __shared__ T Y[NB * RB];
    __shared__ int s_t;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_t = (int)atomicAdd(Y_.ticket, 1u);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= count) break;
        const PTask tk = tasks[t];
        const int J = tk.front;
        const Front f = P.fronts[J];
        if (tid == 0) {
            if (f.parent >= 0) wait_geq(Y_.done + f.parent, Y_.nt[f.parent], F.hang);
            acquire_agent();
        }
        __syncthreads();
        if (tk.kind == PT_SMALL) small_bwd(P, F, S, f, Ps, X);
        else diag_bwd(P, F, S, f, J, tk.a, tk.b, Y_, Ps, X, Y);
        drain_stores();
        __syncthreads();
        if (tid == 0) {
            release_agent();
            if (tk.kind == PT_DIAG) {
                const int p = (f.nn + NB - 1) / NB;
                __hip_atomic_store(Y_.prog + J, p - tk.a / NB, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __hip_atomic_fetch_add(Y_.done + J, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // closes the thread-0 block before the back-edge: every thread of the
        // workgroup reaches the next ticket together
        __syncthreads();
    }
}

template <typename T>
__global__ void k_dscale(const T* d, T* B, int64_t ldb, int n, int nrhs) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * nrhs) return;
    int r = (int)(i % n), c = (int)(i / n);
    B[r + c * ldb] /= d[r];
}

// dst[i] = src[perm[i]] (P*B, multiply.jl:277-280) or dst[perm[i]] = src[i].
template <typename TB>
__global__ void k_permute(const double* src, int64_t lds, TB* dst, int64_t ldd, const int64_t* perm,
                          int n, int nrhs, int inverse) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * nrhs) return;
    int r = (int)(i % n), c = (int)(i / n);
    int64_t p = perm[r] - 1;
    if (!inverse) dst[r + c * ldd] = (TB)src[p + c * lds];
    else dst[p + c * ldd] = (TB)src[r + c * lds];
}

template <typename TS>
__global__ void k_permute_back(const TS* src, int64_t lds, double* dst, int64_t ldd,
                               const int64_t* perm, int n, int nrhs) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * nrhs) return;
    int r = (int)(i % n), c = (int)(i / n);
    dst[(perm[r] - 1) + c * ldd] = (double)src[r + c * lds];
}

inline unsigned blocks_for(int64_t tot) { return (unsigned)((tot + 255) / 256); }

}  // namespace

template <typename T>
void launch_solve(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, int nJ,
                  int32_t* done, int32_t* prog, const int32_t* ntf, const int32_t* ntb,
                  const PTask* ftask, int nf, const PTask* btask, int nb, int n, hipStream_t s) {
    if (nf <= 0) return;
    Sync Y{done, prog, ntf, F.ticket};
    const int grid_f = nf < 1024 ? nf : 1024;
    launch_fill<int32_t>(done, nJ, 0, s);
    launch_fill<int32_t>(prog, nJ, 0, s);
    launch_fill<uint32_t>(F.ticket, 1, 0u, s);
    hipLaunchKernelGGL(k_fsolve<T>, dim3(grid_f), dim3(256), 0, s, P, F, S, ftask, nf, Y);
    if (F.diag != 0 && (int64_t)n * S.nrhs > 0)
        hipLaunchKernelGGL(k_dscale<T>, dim3(blocks_for((int64_t)n * S.nrhs)), dim3(256), 0, s,
                           (const T*)F.d, S.B, S.ldb, n, S.nrhs);
    Y.nt = ntb;
    const int grid_b = nb < 1024 ? nb : 1024;
    launch_fill<int32_t>(done, nJ, 0, s);
    launch_fill<int32_t>(prog, nJ, 0, s);
    launch_fill<uint32_t>(F.ticket, 1, 0u, s);
    hipLaunchKernelGGL(k_bsolve<T>, dim3(grid_b), dim3(256), 0, s, P, F, S, btask, nb, Y);
}

template <typename TB>
void launch_permute(const double* src, int64_t lds, TB* dst, int64_t ldd, const int64_t* perm,
                    int n, int nrhs, bool inverse, hipStream_t s) {
    int64_t tot = (int64_t)n * nrhs;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_permute<TB>, dim3(blocks_for(tot)), dim3(256), 0, s, src, lds, dst, ldd,
                       perm, n, nrhs, inverse ? 1 : 0);
}

void launch_permute_back(const void* src, int64_t lds, bool src_f32, double* dst, int64_t ldd,
                         const int64_t* perm, int n, int nrhs, hipStream_t s) {
    int64_t tot = (int64_t)n * nrhs;
    if (tot <= 0) return;
    if (src_f32)
        hipLaunchKernelGGL(k_permute_back<float>, dim3(blocks_for(tot)), dim3(256), 0, s,
                           (const float*)src, lds, dst, ldd, perm, n, nrhs);
    else
        hipLaunchKernelGGL(k_permute_back<double>, dim3(blocks_for(tot)), dim3(256), 0, s,
                           (const double*)src, lds, dst, ldd, perm, n, nrhs);
}

#define MF_SINST(T)                                                                           \
    template void launch_solve<T>(const DevPlan&, const DevFactor<T>&, const DevSolve<T>&, int, \
                                  int32_t*, int32_t*, const int32_t*, const int32_t*,         \
                                  const PTask*, int, const PTask*, int, int, hipStream_t);    \
    template void launch_permute<T>(const double*, int64_t, T*, int64_t, const int64_t*, int,  \
                                    int, bool, hipStream_t);
MF_SINST(double)
MF_SINST(float)

}  // namespace mf
