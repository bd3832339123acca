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
// and per-front block progress counters.  Thread 0 polls with relaxed loads;
// the data those counters guard is only read and written with
// device-coherent accesses (ldc/stc below), so a task is published by
// draining every wave's stores, a barrier and one relaxed counter update.
// All waits are bounded.  Right-hand sides are processed 16 at a time.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mf_kernels.h"
#include "mf_sync.h"
#include "mf_mfma.h"

namespace mf {

namespace {

constexpr int RB = 16;  // right-hand sides per pass
constexpr int LDS = SMALL_NJ + 1;
constexpr int LDX = NB + 1;  // leading dimension of the 64 x 16 LDS tiles

// Waits until this wave's global stores (and loads) have completed.  Every
// wave drains before the closing barrier of a task, so that the counter
// update thread 0 makes after the barrier follows all of the task's stores.
__device__ __forceinline__ void drain_stores() { __builtin_amdgcn_s_waitcnt(0); }
// Right-hand-side and update-vector entries that other workgroups of the
// same launch read or write go through device-coherent (sc1) loads and
// stores, which the caches keep coherent per access.  Publishing a task
// then needs only its stores to have completed (every wave drains before
// the closing barrier) and a relaxed flag store; no whole-cache L2
// write-back or invalidate, which would also evict the factor tiles that
// other workgroups are reading.  Data written by earlier launches (the
// factor, W, the permuted input) is read with ordinary cached loads.
template <typename T>
__device__ __forceinline__ T ldc(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void stc(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded waits (the watchdog of mf_sync.h; progress is a change of the
// awaited counter or a ticket taken anywhere in the launch).
__device__ __forceinline__ void wait_geq(int32_t* p, int32_t target, int32_t* hang, const uint32_t* ticket,
                                         unsigned long long stall) {
    (void)bounded_wait(p, target, hang, ticket, stall, 2);
}

// Same, returning the value seen (>= target unless the wait gave up), so a
// caller that runs behind the counter does not poll again for every step.
__device__ __forceinline__ int wait_val(int32_t* p, int32_t target, int32_t* hang, const uint32_t* ticket,
                                        unsigned long long stall) {
    return bounded_wait(p, target, hang, ticket, stall, 2);
}

struct Sync {
    int32_t* done;    // per front: finished tasks
    int32_t* prog;    // per front: finished diagonal blocks (in chain order)
    const int32_t* nt;
    uint32_t* ticket;
    unsigned long long* trace;  // nullptr or per ticket: start | workgroup << 48, end
};

// Children's update vectors scattered into Y (rows [r0, r1) of front f, one
// row block: [64b, 64b + 64) or [nn + 64m, ...); column c of the pass at
// Y[(row - r0) + c * LDX]).
template <typename T>
__device__ void gather_children(const DevPlan& P, const DevSolve<T>& S, const Front& f, int r0,
                                int r1, int cb, int nb, T* Y) {
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    const int p = (f.nn + NB - 1) / NB, blk = r0 < f.nn ? r0 / NB : p + (r0 - f.nn) / NB;
    (void)r1;
    for (int e = tid; e < LDX * RB; e += blockDim.x) Y[e] = (T)0;
    __syncthreads();
    for (int q = f.chd0; q < f.chd1; ++q) {
        const int K = P.chd[q];
        const Front kf = P.fronts[K];
        const int ma = kf.na;
        const int32_t* relK = P.rel + kf.reloff;
        // rows of K landing in [r0, r1) (one row block of f): precomputed bounds
        const int32_t* tb = P.sgstart + P.sgoff[K];
        const int a = tb[blk], lo = tb[blk + 1];
        const T* MvK = S.Mv + P.mvoff1[K] * S.nrhs;
        for (int i = a + r; i < lo; i += 64) {
            const int yr = relK[i] - r0;
            for (int c = g; c < nb; c += 4) Y[yr + c * LDX] += ldc(MvK + i + (int64_t)(cb + c) * ma);
        }
        __syncthreads();
    }
}

// ---- 64 x 64 tile products of the big fronts on the matrix cores --------------
// A wave owns 16 rows of a 64-row block of right-hand sides; lane (li, lk)
// holds the A operand A[16w + li][4s + lk] for s = 0..15 in registers (so
// the factor tile of the next step is loaded while the current step waits),
// the B operand (64 rows x 16 right-hand sides) is staged in LDS.

// a[s] = sgn * base[i * si + j * sj] for i = 16w + li < rows, j = 4s + lk < cols, else 0.
template <typename T>
__device__ __forceinline__ void load_frag(const T* base, int64_t si, int64_t sj, int rows, int cols,
                                          T sgn, T (&a)[NB / 4]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int i = 16 * (threadIdx.x >> 6) + li;
    // one running pointer (the column stride is not a compile-time constant)
    const T* q = base + (int64_t)i * si + (int64_t)lk * sj;
    const int64_t step = 4 * sj;
    const bool row_ok = i < rows;
    // every load unconditional (base, element (0, 0), stands in outside the
    // tile) and the predicate and sign applied afterwards: a load whose value
    // is used inside a per-lane branch is waited for inside that branch
    T v[NB / 4];
#pragma unroll
    for (int s = 0; s < NB / 4; ++s) {
        v[s] = (row_ok && 4 * s + lk < cols) ? *q : *base;
        q += step;
    }
#pragma unroll
    for (int s = 0; s < NB / 4; ++s) a[s] = (row_ok && 4 * s + lk < cols) ? sgn * v[s] : (T)0;
}

// acc += A (registers) * Xs (64 x 16, LDS, leading dimension LDX)
template <typename T>
__device__ __forceinline__ void frag_mma(const T (&a)[NB / 4], const T* Xs,
                                         typename Mfma<T>::v4& acc) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int s = 0; s < NB / 4; ++s) acc = Mfma<T>::op(a[s], Xs[(4 * s + lk) + li * LDX], acc);
}

// Row of output register v of this lane inside the 64-row block.
template <typename T>
__device__ __forceinline__ int acc_row(int v) {
    return 16 * (threadIdx.x >> 6) + Mfma<T>::row((threadIdx.x & 63) >> 4, v);
}

// A 64-row block of right-hand sides in registers, thread (r, g) holding
// B(row0 + r, cb + g + 4i) (device-coherent loads; 0 outside rows x nb), and
// its copy into an LDS tile Xs (leading dimension LDX).  Fetching the next
// block before the current product hides its latency.
template <typename T>
__device__ __forceinline__ void fetch_rows(const T* B, int64_t ldb, int row0, int rows, int cb, int nb,
                                           T (&x)[RB / 4]) {
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < RB / 4; ++i) {
        const int c = g + 4 * i;
        x[i] = (r < rows && c < nb) ? ldc(B + row0 + r + (int64_t)(cb + c) * ldb) : (T)0;
    }
}
template <typename T>
__device__ __forceinline__ void fetch_gather(const T* B, int64_t ldb, const int32_t* idx, int rows, int cb,
                                             int nb, T (&x)[RB / 4]) {
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < RB / 4; ++i) {
        const int c = g + 4 * i;
        x[i] = (r < rows && c < nb) ? ldc(B + idx[r] + (int64_t)(cb + c) * ldb) : (T)0;
    }
}
template <typename T>
__device__ __forceinline__ void put_rows(const T (&x)[RB / 4], T* Xs) {
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < RB / 4; ++i) Xs[r + (g + 4 * i) * LDX] = x[i];
}

// X_k = W acc for the block's rows < kb: acc goes through LDS (Ys) as the B
// operand; W's fragment is already in registers.
template <typename T>
__device__ __forceinline__ typename Mfma<T>::v4 times_w(const T (&wa)[NB / 4], typename Mfma<T>::v4 acc,
                                                        T* Ys) {
    typedef typename Mfma<T>::v4 v4;
    const int li = threadIdx.x & 15;
    __syncthreads();  // every wave is done reading Ys / Xs of the previous step
#pragma unroll
    for (int v = 0; v < 4; ++v) Ys[acc_row<T>(v) + li * LDX] = acc[v];
    __syncthreads();
    v4 x = v4{0, 0, 0, 0};
    frag_mma(wa, Ys, x);
    return x;
}

// Loads of the small-front paths: every load of a loop issued first, from a
// clamped (always valid) address, then the LDS updates.  A load whose value
// is stored or added inside a per-lane branch or loop is otherwise waited
// for at once, one memory round trip per element.

// Ps[r + c * ldp] = Pj[r + c * nj] for c = c0, c0 + dc, ... < nn (rows r < nj)
template <typename T, int ldp, int dc>
__device__ __forceinline__ void load_panel(const T* Pj, int nj, int nn, int r, int c0, T* Ps) {
    for (int cb = c0; cb < nn; cb += 8 * dc) {
        T v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int c = cb + dc * i;
            v[i] = Pj[(r < nj && c < nn) ? r + (int64_t)c * nj : 0];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int c = cb + dc * i;
            if (r < nj && c < nn) Ps[r + c * ldp] = v[i];
        }
    }
}

// X[xr + c * ldx] (+)= src[c * lds] for c = c0, c0 + dc, ... < nb (RB / dc values)
template <typename T, int dc, bool add, bool coherent>
__device__ __forceinline__ void load_rhs(const T* src, int64_t lds, int nb, int c0, bool on, T* X, int xr,
                                         int ldx) {
#pragma unroll
    for (int h = 0; h < RB / dc; h += 8) {
        constexpr int M = RB / dc < 8 ? RB / dc : 8;
        T v[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int c = c0 + dc * (h + i);
            const T* a = src + (on && c < nb ? c * lds : 0);
            v[i] = coherent ? ldc(a) : *a;
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int c = c0 + dc * (h + i);
            if (on && c < nb) {
                if (add) X[xr + c * ldx] += v[i];
                else X[xr + c * ldx] = v[i];
            }
        }
    }
}

template <typename T>
__device__ void small_fwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                          const Front& f, int J, T* Ps, T* X) {
    const int tid = threadIdx.x, r = tid & 63, g = tid >> 6;
    const int nn = f.nn, na = f.na, nj = nn + na;
    const bool unit = F.diag != 0;
    const T* Pj = F.panel + f.poff;
    load_panel<T, LDS, 4>(Pj, nj, nn, r, g, Ps);
    T* MvJ = S.Mv + P.mvoff1[J] * S.nrhs;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        for (int c = g; c < RB; c += 4) X[r + c * LDS] = (T)0;
        __syncthreads();
        for (int q = f.chd0; q < f.chd1; ++q) {
            const int K = P.chd[q];
            const Front kf = P.fronts[K];
            const int ma = kf.na;
            const int rr = P.rel[kf.reloff + (r < ma ? r : 0)];
            const T* MvK = S.Mv + P.mvoff1[K] * S.nrhs;
            load_rhs<T, 4, true, true>(MvK + (r < ma ? r : 0) + (int64_t)cb * ma, ma, nb, g, r < ma, X, rr, LDS);
            __syncthreads();
        }
        load_rhs<T, 4, true, false>(S.B + f.c0 + (r < nn ? r : 0) + (int64_t)cb * S.ldb, S.ldb, nb, g, r < nn, X,
                                    r, LDS);
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
                if (r < nn) stc(S.B + f.c0 + r + (int64_t)(cb + c) * S.ldb, X[r + c * LDS]);
                else stc(MvJ + (r - nn) + (int64_t)(cb + c) * na, X[r + c * LDS]);
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
    load_panel<T, LDS, 4>(Pj, nj, nn, r, g, Ps);
    const int32_t* rows = P.rows + f.rowoff;
    const int rowr = rows[r < nj ? r : 0];
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        load_rhs<T, 4, false, true>(S.B + rowr + (int64_t)cb * S.ldb, S.ldb, nb, g, r < nj, X, r, LDS);
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
            for (int c = g; c < nb; c += 4) stc(S.B + f.c0 + r + (int64_t)(cb + c) * S.ldb, X[r + c * LDS]);
        __syncthreads();
    }
}

// Tiny fronts (nJ <= TINY_NJ), one wavefront each: the panel in a private
// LDS slab Pw (leading dimension LDT1), the right-hand sides of the pass in
// Xw (leading dimension TINY_NJ); lane r keeps row r.  A wavefront's LDS
// operations complete in order, so no barrier is needed.
constexpr int LDT1 = TINY_NJ + 1;

template <typename T>
__device__ void wave_fwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, int J, T* Pw,
                         T* Xw) {
    const int r = threadIdx.x & 63;
    const Front f = P.fronts[J];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const bool unit = F.diag != 0;
    const T* Pj = F.panel + f.poff;
    load_panel<T, LDT1, 1>(Pj, nj, nn, r, 0, Pw);
    T* MvJ = S.Mv + P.mvoff1[J] * S.nrhs;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        if (r < TINY_NJ)
            for (int c = 0; c < RB; ++c) Xw[r + c * TINY_NJ] = (T)0;
        for (int q = f.chd0; q < f.chd1; ++q) {
            const int K = P.chd[q];
            const Front kf = P.fronts[K];
            const int ma = kf.na;
            const int rr = P.rel[kf.reloff + (r < ma ? r : 0)];
            const T* MvK = S.Mv + P.mvoff1[K] * S.nrhs;
            load_rhs<T, 1, true, true>(MvK + (r < ma ? r : 0) + (int64_t)cb * ma, ma, nb, 0, r < ma, Xw, rr,
                                       TINY_NJ);
        }
        load_rhs<T, 1, true, false>(S.B + f.c0 + (r < nn ? r : 0) + (int64_t)cb * S.ldb, S.ldb, nb, 0, r < nn, Xw,
                                    r, TINY_NJ);
        for (int k = 0; k < nn; ++k) {
            if (!unit && r < nb) Xw[k + r * TINY_NJ] /= Pw[k + k * LDT1];
            if (r > k && r < nj) {
                const T p = Pw[r + k * LDT1];
                for (int c = 0; c < nb; ++c) Xw[r + c * TINY_NJ] -= p * Xw[k + c * TINY_NJ];
            }
        }
        if (r < nj) {
            for (int c = 0; c < nb; ++c) {
                if (r < nn) stc(S.B + f.c0 + r + (int64_t)(cb + c) * S.ldb, Xw[r + c * TINY_NJ]);
                else stc(MvJ + (r - nn) + (int64_t)(cb + c) * na, Xw[r + c * TINY_NJ]);
            }
        }
    }
}

template <typename T>
__device__ void wave_bwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, int J, T* Pw,
                         T* Xw) {
    const int r = threadIdx.x & 63;
    const Front f = P.fronts[J];
    const int nn = f.nn, nj = nn + f.na;
    const bool unit = F.diag != 0;
    const T* Pj = F.panel + f.poff;
    load_panel<T, LDT1, 1>(Pj, nj, nn, r, 0, Pw);
    const int32_t* rows = P.rows + f.rowoff;
    const int rowr = rows[r < nj ? r : 0];
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        load_rhs<T, 1, false, true>(S.B + rowr + (int64_t)cb * S.ldb, S.ldb, nb, 0, r < nj, Xw, r, TINY_NJ);
        if (r < nn) {
            for (int c = 0; c < nb; ++c) {
                T s = (T)0;
                for (int q = nn; q < nj; ++q) s += Pw[q + r * LDT1] * Xw[q + c * TINY_NJ];
                Xw[r + c * TINY_NJ] -= s;
            }
        }
        for (int k = nn - 1; k >= 0; --k) {
            if (!unit && r < nb) Xw[k + r * TINY_NJ] /= Pw[k + k * LDT1];
            if (r < k) {
                const T p = Pw[k + r * LDT1];
                for (int c = 0; c < nb; ++c) Xw[r + c * TINY_NJ] -= p * Xw[k + c * TINY_NJ];
            }
        }
        if (r < nn)
            for (int c = 0; c < nb; ++c) stc(S.B + f.c0 + r + (int64_t)(cb + c) * S.ldb, Xw[r + c * TINY_NJ]);
    }
}

// Forward diagonal block [k0, k0+kb) of a big front: acc = B_k + gathered
// children - sum_{ib<k} L_{k,ib} X_ib, taking X_ib as the front's progress
// counter releases them, then X_k = W_k acc.
template <typename T>
__device__ void diag_fwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                         const Front& f, int J, int k0, int kb, const Sync& Y_, T* X, T* Y, int* s_av) {
    typedef typename Mfma<T>::v4 v4;
    const int tid = threadIdx.x, li = tid & 15;
    const int nj = f.nn + f.na;
    const T* Pj = F.panel + f.poff;
    const T* W = F.W + f.woff;
    const int kblk = k0 / NB;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        T wa[NB / 4], la[NB / 4];
        load_frag(W + k0, 1, f.nn, kb, kb, (T)1, wa);
        if (kblk > 0) load_frag(Pj + k0, 1, nj, kb, NB, (T)-1, la);
        gather_children(P, S, f, k0, k0 + kb, cb, nb, Y);
        v4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = acc_row<T>(v);
            acc[v] = (row < kb && li < nb)
                         ? S.B[f.c0 + k0 + row + (int64_t)(cb + li) * S.ldb] + Y[row + li * LDX]
                         : (T)0;
        }
        // thread 0 keeps the last progress it saw and polls only when it has
        // caught up; a block already published is fetched one step ahead
        int avail = 0;
        bool have = false;
        T xr[RB / 4];
        for (int ib = 0; ib < kblk; ++ib) {
            if (tid == 0) {
                if (ib >= avail) avail = wait_val(Y_.prog + J, ib + 1, F.hang, F.ticket, F.stall);
                *s_av = avail;
            }
            __syncthreads();
            const int av = *s_av;
            if (!have) fetch_rows(S.B, S.ldb, f.c0 + ib * NB, NB, cb, nb, xr);
            put_rows(xr, X);
            have = ib + 1 < kblk && ib + 1 < av;
            if (have) fetch_rows(S.B, S.ldb, f.c0 + (ib + 1) * NB, NB, cb, nb, xr);
            __syncthreads();
            frag_mma(la, X, acc);
            if (ib + 1 < kblk) load_frag(Pj + k0 + (int64_t)(ib + 1) * NB * nj, 1, nj, kb, NB, (T)-1, la);
        }
        const v4 x = times_w(wa, acc, Y);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = acc_row<T>(v);
            if (row < kb && li < nb) stc(S.B + f.c0 + k0 + row + (int64_t)(cb + li) * S.ldb, x[v]);
        }
        __syncthreads();
    }
}

// Forward update rows [r0, r1) (>= nn) of a big front: M2 = gather - L21 X.
template <typename T>
__device__ void rows_fwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                         const Front& f, int J, int r0, int r1, const Sync& Y_, T* X, T* Y, int* s_av) {
    typedef typename Mfma<T>::v4 v4;
    const int tid = threadIdx.x, li = tid & 15;
    const int nn = f.nn, na = f.na, nj = nn + na, m = r1 - r0;
    const T* Pj = F.panel + f.poff;
    const int p = (nn + NB - 1) / NB;
    T* MvJ = S.Mv + P.mvoff1[J] * S.nrhs;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        T la[NB / 4];
        load_frag(Pj + r0, 1, nj, m, min(NB, nn), (T)-1, la);
        gather_children(P, S, f, r0, r1, cb, nb, Y);
        v4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = acc_row<T>(v);
            acc[v] = (row < m && li < nb) ? Y[row + li * LDX] : (T)0;
        }
        int avail = 0;
        bool have = false;
        T xr[RB / 4];
        for (int ib = 0; ib < p; ++ib) {
            const int kbi = min(NB, nn - ib * NB);
            if (tid == 0) {
                if (ib >= avail) avail = wait_val(Y_.prog + J, ib + 1, F.hang, F.ticket, F.stall);
                *s_av = avail;
            }
            __syncthreads();
            const int av = *s_av;
            if (!have) fetch_rows(S.B, S.ldb, f.c0 + ib * NB, kbi, cb, nb, xr);
            put_rows(xr, X);
            have = ib + 1 < p && ib + 1 < av;
            if (have) fetch_rows(S.B, S.ldb, f.c0 + (ib + 1) * NB, min(NB, nn - (ib + 1) * NB), cb, nb, xr);
            __syncthreads();
            frag_mma(la, X, acc);
            if (ib + 1 < p)
                load_frag(Pj + r0 + (int64_t)(ib + 1) * NB * nj, 1, nj, m, min(NB, nn - (ib + 1) * NB),
                          (T)-1, la);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = acc_row<T>(v);
            if (row < m && li < nb) stc(MvJ + (r0 + row - nn) + (int64_t)(cb + li) * na, acc[v]);
        }
        __syncthreads();
    }
}

// Separator part of a backward diagonal block, split over the four
// wavefronts: wave w takes the separator tiles q = w, w + 4, ... (64 rows
// each, as eight chunks of 8) and accumulates part[i] = -sum L21(rows,
// block)' x_sep(rows) for all four 16-row slices i of the block, with no
// barrier between chunks.  The front's separator row indices sit in LDS
// (sidx, local row - nn), so the only global loads in flight are the
// operands: those of chunk u + 1 load while chunk u is multiplied.  Lane
// (li, lk) of chunk u holds rows nn + 64q + 8c + 4s + lk (c = u % 8, s < 2).
constexpr int SEP_S = 2;  // 4-row MFMA steps per chunk

// r0: the first separator row of the tiles taken (a chunk of the
// separator); sidx holds the chunk's row indices (sidx[0] is row r0).
template <typename T>
__device__ __forceinline__ void sep_operands(const T* Pj, int64_t nj, int nn, int k0, int kb, const T* B,
                                             int64_t ldb, int cb, int nb, int wv, int u, const int32_t* sidx,
                                             int r0, T (&a)[4 * SEP_S], T (&b)[SEP_S]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int rel = 64 * (wv + 4 * (u >> 3)) + 8 * (u & 7) + lk;  // row in the chunk
    const int base = r0 + rel;                                      // separator row (local - nn)
    const int na = (int)nj - nn;
#pragma unroll
    for (int s = 0; s < SEP_S; ++s)
        b[s] = (base + 4 * s < na && li < nb) ? ldc(B + sidx[rel + 4 * s] + (int64_t)(cb + li) * ldb) : (T)0;
    // loads unconditional (the block's first separator entry stands in), the
    // predicate and sign afterwards, so that none is waited for on its own
    const T* anchor = Pj + (int64_t)k0 * nj + nn;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = 16 * i + li;
        const T* col = Pj + (int64_t)(k0 + m) * nj + nn + base;
#pragma unroll
        for (int s = 0; s < SEP_S; ++s) a[SEP_S * i + s] = (m < kb && base + 4 * s < na) ? col[4 * s] : *anchor;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int s = 0; s < SEP_S; ++s)
            a[SEP_S * i + s] = (16 * i + li < kb && base + 4 * s < na) ? -a[SEP_S * i + s] : (T)0;
}
template <typename T>
__device__ __forceinline__ void sep_partial(const T* Pj, int nj, int nn, int nsep, int k0, int kb,
                                            const int32_t* sidx, const T* B, int64_t ldb, int cb, int nb,
                                            typename Mfma<T>::v4 (&part)[4], int r0 = 0) {
    const int wv = threadIdx.x >> 6;
    const int nu = nsep > wv ? 8 * ((nsep - wv + 3) / 4) : 0;
    if (nu == 0) return;
    T a0[4 * SEP_S], b0[SEP_S], a1[4 * SEP_S], b1[SEP_S];
    auto mma = [&](const T (&a)[4 * SEP_S], const T (&b)[SEP_S]) {
#pragma unroll
        for (int s = 0; s < SEP_S; ++s)
#pragma unroll
            for (int i = 0; i < 4; ++i) part[i] = Mfma<T>::op(a[SEP_S * i + s], b[s], part[i]);
    };
    sep_operands(Pj, nj, nn, k0, kb, B, ldb, cb, nb, wv, 0, sidx, r0, a0, b0);
    // nu is a multiple of 8: chunks alternate between the two operand sets
    for (int u = 0; u < nu; u += 2) {
        sep_operands(Pj, nj, nn, k0, kb, B, ldb, cb, nb, wv, u + 1, sidx, r0, a1, b1);
        mma(a0, b0);
        if (u + 2 < nu) sep_operands(Pj, nj, nn, k0, kb, B, ldb, cb, nb, wv, u + 2, sidx, r0, a0, b0);
        mma(a1, b1);
    }
}

// LDS elements of a solve workgroup's panel buffer (a small front, four tiny
// fronts, or the separator indices and partial sums of diag_bwd)
constexpr int NPS = SMALL_NJ * LDS > 4 * TINY_NJ * LDT1 ? SMALL_NJ * LDS : 4 * TINY_NJ * LDT1;

// Backward separator product of diagonal block [k0, k0 + kb) over the
// separator tiles [t0, t1) (PT_SEPB): part = -L21(chunk rows, block)'
// x_sep(chunk rows), the four wavefronts' sums added in wave order, stored to
// the block's partial slot for the chunk (64 x nrhs, column stride 64).  The
// front's diagonal tasks add the slots in chunk order, so the result does
// not depend on timing.
template <typename T>
__device__ __attribute__((noinline)) void sep_bwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, const Front& f,
                        int J, int k0, int t0, int t1, T* Xbuf, T* Red) {
    typedef typename Mfma<T>::v4 v4;
    const int tid = threadIdx.x, li = tid & 15, lk = (tid & 63) >> 4, wv = tid >> 6;
    const int nn = f.nn, nj = nn + f.na, kb = min(NB, nn - k0);
    const T* Pj = F.panel + f.poff;
    const int nsep = (f.na + NB - 1) / NB, nchk = (nsep + SEP_TILES - 1) / SEP_TILES;
    T* slot = S.Mv + (P.sboff1[J] + (int64_t)((k0 / NB) * nchk + t0 / SEP_TILES) * NB) * S.nrhs;
    // the chunk's separator row indices in LDS (the X/Y buffer is unused here)
    const int r0 = NB * t0, nr = min(f.na, NB * t1) - r0;
    int32_t* si = reinterpret_cast<int32_t*>(Xbuf);
    const int32_t* rows = P.rows + f.rowoff + nn + r0;
    for (int e = tid; e < nr; e += blockDim.x) si[e] = rows[e];
    __syncthreads();
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        v4 part[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) part[i] = v4{0, 0, 0, 0};
        sep_partial(Pj, nj, nn, t1 - t0, k0, kb, si, S.B, S.ldb, cb, nb, part, r0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) Red[wv * 1024 + 16 * i + Mfma<T>::row(lk, v) + li * 64] = part[i][v];
        __syncthreads();
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = acc_row<T>(v), e = row + li * 64;
            const T x = ((Red[e] + Red[1024 + e]) + Red[2048 + e]) + Red[3072 + e];
            if (row < kb && li < nb) stc(slot + row + (int64_t)(cb + li) * NB, x);
        }
        __syncthreads();
    }
}

// Backward diagonal block [k0, k0+kb): acc = C_k - L21' x_sep - sum_{ib>k}
// L_{ib,k}' X_ib (x_sep final: the ancestors are complete; X_ib as the
// progress counter releases them, last block first), then X_k = W_k' acc.
// The separator part is split over the wavefronts (sep_partial) and summed
// in wave order through Red (4 x 64 x 16 elements of LDS), so the result
// does not depend on timing.
template <typename T>
__device__ void diag_bwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                         const Front& f, int J, int k0, int kb, const Sync& Y_, T* X, T* Y, T* Red,
                         int* s_av) {
    typedef typename Mfma<T>::v4 v4;
    const int tid = threadIdx.x, li = tid & 15, lk = (tid & 63) >> 4, wv = tid >> 6;
    const int nn = f.nn, nj = nn + f.na;
    const T* Pj = F.panel + f.poff;
    const T* W = F.W + f.woff;
    const int32_t* rows = P.rows + f.rowoff;
    const int p = (nn + NB - 1) / NB, kblk = k0 / NB;
    const int nsep = (nj - nn + NB - 1) / NB, nch = p - 1 - kblk;
    const bool split = P.sboff1[J] >= 0;
    const int nchk = (nsep + SEP_TILES - 1) / SEP_TILES;
    if (split) {
        // every PT_SEPB task of the front is done: no diagonal task of the
        // front can finish before that, so done[J] counts only them so far
        if (tid == 0) wait_geq(Y_.done + J, p * nchk, F.hang, F.ticket, F.stall);
        __syncthreads();
    }
    // A fragment of chain step t: L_{ib,k}' for ib = p - 1 - t
    auto cfrag = [&](int t, T (&a)[NB / 4]) {
        const int ib = p - 1 - t;
        load_frag(Pj + ib * NB + (int64_t)k0 * nj, nj, 1, kb, min(NB, nn - ib * NB), (T)-1, a);
    };
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        v4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = acc_row<T>(v);
            acc[v] = (row < kb && li < nb) ? S.B[f.c0 + k0 + row + (int64_t)(cb + li) * S.ldb] : (T)0;
        }
        if (split) {
            // the separator product computed by the front's PT_SEPB tasks:
            // their partials (64 x nrhs per chunk), summed in chunk order
            const T* slot0 = S.Mv + (P.sboff1[J] + (int64_t)kblk * nchk * NB) * S.nrhs;
            for (int c = 0; c < nchk; ++c) {
                const T* sl = slot0 + (int64_t)c * NB * S.nrhs;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = acc_row<T>(v);
                    const T x = ldc(sl + min(row, kb - 1) + (int64_t)min(cb + li, S.nrhs - 1) * NB);
                    acc[v] += (row < kb && li < nb) ? x : (T)0;
                }
            }
        } else if (nsep > 0) {
            v4 part[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) part[i] = v4{0, 0, 0, 0};
            // separator row indices in LDS (Red's space until the reduction);
            // fronts with more separator rows than fit read them from rows[]
            const int32_t* sidx = rows + nn;
            if (nj - nn <= NPS * (int)sizeof(T) / 4) {
                int32_t* si = reinterpret_cast<int32_t*>(Red);
                for (int e = tid; e < nj - nn; e += blockDim.x) si[e] = rows[nn + e];
                __syncthreads();
                sidx = si;
            }
            sep_partial(Pj, nj, nn, nsep, k0, kb, sidx, S.B, S.ldb, cb, nb, part);
            __syncthreads();  // every wave is done with the indices in Red
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) Red[wv * 1024 + 16 * i + Mfma<T>::row(lk, v) + li * 64] = part[i][v];
            __syncthreads();
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int e = acc_row<T>(v) + li * 64;
                acc[v] += ((Red[e] + Red[1024 + e]) + Red[2048 + e]) + Red[3072 + e];
            }
        }
        // W and the first chain operand after the separator part (fewer live
        // registers during it)
        T wa[NB / 4], la[NB / 4];
        load_frag(W + k0, f.nn, 1, kb, kb, (T)1, wa);
        if (nch > 0) cfrag(0, la);
        int avail = 0;
        bool have = false;
        T xr[RB / 4];
        for (int t = 0; t < nch; ++t) {
            if (tid == 0) {
                if (t + 1 > avail) avail = wait_val(Y_.prog + J, t + 1, F.hang, F.ticket, F.stall);
                *s_av = avail;
            }
            __syncthreads();
            const int av = *s_av;
            const int ib = p - 1 - t;
            if (!have) fetch_rows(S.B, S.ldb, f.c0 + ib * NB, min(NB, nn - ib * NB), cb, nb, xr);
            put_rows(xr, X);
            have = t + 1 < nch && t + 2 <= av;
            if (have) fetch_rows(S.B, S.ldb, f.c0 + (ib - 1) * NB, min(NB, nn - (ib - 1) * NB), cb, nb, xr);
            __syncthreads();
            frag_mma(la, X, acc);
            if (t + 1 < nch) cfrag(t + 1, la);
        }
        const v4 x = times_w(wa, acc, Y);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = acc_row<T>(v);
            if (row < kb && li < nb) stc(S.B + f.c0 + k0 + row + (int64_t)(cb + li) * S.ldb, x[v]);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256, 2) void k_fsolve(DevPlan P, DevFactor<T> F, DevSolve<T> S,
                                                const PTask* tasks, int count, Sync Y_) {
    static_assert(2 * LDX * RB >= 4 * TINY_NJ * RB, "tiny right-hand-side slabs");
    __shared__ T Ps[NPS];
    __shared__ T XY[2 * LDX * RB];  // X and Y of the big-front tasks, or four tiny slabs
    T* const X = XY;
    T* const Y = XY + LDX * RB;
    __shared__ int s_t, s_av;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_t = (int)atomicAdd(Y_.ticket, 1u);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= count) break;
        const unsigned long long t_begin = (Y_.trace && tid == 0) ? wall_clock64() : 0;
        const PTask tk = tasks[t];
        if (tk.kind == PT_TINY) {
            const int w = tid >> 6, lane = tid & 63;
            const int J = w == 0 ? tk.front : w == 1 ? tk.a : w == 2 ? tk.b : tk.c;
            if (J >= 0) {
                const Front f = P.fronts[J];
                if (lane == 0)
                    for (int q = f.chd0; q < f.chd1; ++q) {
                        const int K = P.chd[q];
                        wait_geq(Y_.done + K, Y_.nt[K], F.hang, F.ticket, F.stall);
                    }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                wave_fwd(P, F, S, J, Ps + w * TINY_NJ * LDT1, XY + w * TINY_NJ * RB);
                drain_stores();
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) __hip_atomic_fetch_add(Y_.done + J, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (Y_.trace && tid == 0) {
                Y_.trace[2 * t] = t_begin | ((unsigned long long)blockIdx.x << 48);
                Y_.trace[2 * t + 1] = wall_clock64();
            }
            __syncthreads();
            continue;
        }
        const int J = tk.front;
        const Front f = P.fronts[J];
        if (tid == 0) {
            for (int q = f.chd0; q < f.chd1; ++q) {
                const int K = P.chd[q];
                wait_geq(Y_.done + K, Y_.nt[K], F.hang, F.ticket, F.stall);
            }
        }
        __syncthreads();
        if (tk.kind == PT_SMALL) small_fwd(P, F, S, f, J, Ps, X);
        else if (tk.kind == PT_DIAG) diag_fwd(P, F, S, f, J, tk.a, tk.b, Y_, X, Y, &s_av);
        else rows_fwd(P, F, S, f, J, tk.a, tk.b, Y_, X, Y, &s_av);
        drain_stores();
        __syncthreads();
        if (tid == 0) {
            if (tk.kind == PT_DIAG)
                __hip_atomic_store(Y_.prog + J, tk.a / NB + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(Y_.done + J, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (Y_.trace) {
                Y_.trace[2 * t] = t_begin | ((unsigned long long)blockIdx.x << 48);
                Y_.trace[2 * t + 1] = wall_clock64();
            }
        }
        // closes the thread-0 block before the back-edge: every thread of the
        // workgroup reaches the next ticket together
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256, 2) void k_bsolve(DevPlan P, DevFactor<T> F, DevSolve<T> S,
                                                const PTask* tasks, int count, Sync Y_) {
    static_assert(2 * LDX * RB >= 4 * TINY_NJ * RB, "tiny right-hand-side slabs");
    static_assert(NPS >= 4 * NB * RB, "separator partial sums of diag_bwd in Ps");
    __shared__ T Ps[NPS];
    __shared__ T XY[2 * LDX * RB];  // X and Y of the big-front tasks, or four tiny slabs
    T* const X = XY;
    T* const Y = XY + LDX * RB;
    __shared__ int s_t, s_av;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_t = (int)atomicAdd(Y_.ticket, 1u);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= count) break;
        const unsigned long long t_begin = (Y_.trace && tid == 0) ? wall_clock64() : 0;
        const PTask tk = tasks[t];
        if (tk.kind == PT_TINY) {
            const int w = tid >> 6, lane = tid & 63;
            const int J = w == 0 ? tk.front : w == 1 ? tk.a : w == 2 ? tk.b : tk.c;
            if (J >= 0) {
                const Front f = P.fronts[J];
                if (lane == 0 && f.parent >= 0) wait_geq(Y_.done + f.parent, Y_.nt[f.parent], F.hang, F.ticket, F.stall);
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                wave_bwd(P, F, S, J, Ps + w * TINY_NJ * LDT1, XY + w * TINY_NJ * RB);
                drain_stores();
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) __hip_atomic_fetch_add(Y_.done + J, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (Y_.trace && tid == 0) {
                Y_.trace[2 * t] = t_begin | ((unsigned long long)blockIdx.x << 48);
                Y_.trace[2 * t + 1] = wall_clock64();
            }
            __syncthreads();
            continue;
        }
        const int J = tk.front;
        const Front f = P.fronts[J];
        if (tid == 0) {
            if (f.parent >= 0) wait_geq(Y_.done + f.parent, Y_.nt[f.parent], F.hang, F.ticket, F.stall);
        }
        __syncthreads();
        if (tk.kind == PT_SMALL) small_bwd(P, F, S, f, Ps, X);
        else if (tk.kind == PT_SEPB) sep_bwd(P, F, S, f, J, tk.a, tk.b, tk.c, XY, Ps);
        else diag_bwd(P, F, S, f, J, tk.a, tk.b, Y_, X, Y, Ps, &s_av);
        drain_stores();
        __syncthreads();
        if (tid == 0) {
            if (tk.kind == PT_DIAG) {
                const int p = (f.nn + NB - 1) / NB;
                __hip_atomic_store(Y_.prog + J, p - tk.a / NB, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __hip_atomic_fetch_add(Y_.done + J, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (Y_.trace) {
                Y_.trace[2 * t] = t_begin | ((unsigned long long)blockIdx.x << 48);
                Y_.trace[2 * t + 1] = wall_clock64();
            }
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
                  const PTask* ftask, int nf, const PTask* btask, int nb, int n,
                  unsigned long long* trace, hipStream_t s) {
    if (nf <= 0) return;
    Sync Y{done, prog, ntf, F.ticket, trace};
    const int grid_f = nf < 1024 ? nf : 1024;
    launch_fill<int32_t>(done, nJ, 0, s);
    launch_fill<int32_t>(prog, nJ, 0, s);
    launch_fill<uint32_t>(F.ticket, 1, 0u, s);
    hipLaunchKernelGGL(k_fsolve<T>, dim3(grid_f), dim3(256), 0, s, P, F, S, ftask, nf, Y);
    if (F.diag != 0 && (int64_t)n * S.nrhs > 0)
        hipLaunchKernelGGL(k_dscale<T>, dim3(blocks_for((int64_t)n * S.nrhs)), dim3(256), 0, s,
                           (const T*)F.d, S.B, S.ldb, n, S.nrhs);
    Y.nt = ntb;
    if (trace) Y.trace = trace + 2 * (size_t)nf;
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
                                  const PTask*, int, const PTask*, int, int,                  \
                                  unsigned long long*, hipStream_t);                          \
    template void launch_permute<T>(const double*, int64_t, T*, int64_t, const int64_t*, int,  \
                                    int, bool, hipStream_t);
MF_SINST(double)
MF_SINST(float)

}  // namespace mf
