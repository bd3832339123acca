// mf_solve.hip — gfx950 kernels of the supernodal triangular solves.
//
// Forward  C <- L^-1 C  (div_impl! forward, divide.jl:417-450; per front:
//   F <- sum_i R_i M_i (div_fwd_update!, divide.jl:1206-1248), C1 += F1,
//   C1 <- D11^-1 C1, M2 <- F2 - L21 C1, divide.jl:493-659).
// Backward C <- L^-T C (divide.jl:864-1038; C1 -= L21' M2 with M2 = the
//   already final rows sep(j) of C, then C1 <- D11^-T C1).
// Small internal fronts (nJ <= 64) run one workgroup per front with the
// panel in LDS (persistent, ticket order, children/parent completion flags).
// Big fronts run level by level: a gather of the children's update vectors,
// then per 64-column panel a diagonal-block product with the inverted block
// (W) and a row-block update.  RHS are processed 16 columns at a time.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mf_kernels.h"

namespace mf {

namespace {

constexpr int RB = 16;             // right-hand sides per pass
constexpr int LDS = SMALL_NJ + 1;

__device__ __forceinline__ void release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

__device__ __forceinline__ void wait_flag(int32_t* flag, int32_t* hang) {
    int spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (++spins > (1 << 24)) {
            atomicExch(hang, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_ssmall_fwd(DevPlan P, DevFactor<T> F, DevSolve<T> S,
                                                    const int32_t* list, int count,
                                                    int persistent) {
    __shared__ T Ps[SMALL_NJ * LDS];
    __shared__ T X[LDS * RB];
    __shared__ int s_t;
    const int tid = threadIdx.x;
    const int unit = F.diag != 0;
    int t = blockIdx.x;
    for (;;) {
        if (persistent) {
            if (tid == 0) s_t = (int)atomicAdd(F.ticket, 1u);
            __syncthreads();
            t = s_t;
        }
        if (t >= count) return;
        const int J = list[t];
        const Front f = P.fronts[J];
        const int nn = f.nn, na = f.na, nj = nn + na;
        if (tid == 0 && persistent)
            for (int q = f.chd0; q < f.chd1; ++q) wait_flag(F.flags + P.chd[q], F.hang);
        __syncthreads();
        acquire_agent();
        const T* Pj = F.panel + f.poff;
        for (int e = tid; e < nj * nn; e += blockDim.x) {
            int r = e % nj, c = e / nj;
            Ps[r + c * LDS] = Pj[e];
        }
        T* MvJ = S.Mv + P.mvoff1[J] * S.nrhs;
        for (int cb = 0; cb < S.nrhs; cb += RB) {
            const int nb = min(RB, S.nrhs - cb);
            for (int e = tid; e < LDS * RB; e += blockDim.x) X[e] = (T)0;
            __syncthreads();
            for (int q = f.chd0; q < f.chd1; ++q) {
                const int K = P.chd[q];
                const Front kf = P.fronts[K];
                const int ma = kf.na;
                const int32_t* relK = P.rel + kf.reloff;
                const T* MvK = S.Mv + P.mvoff1[K] * S.nrhs;
                for (int e = tid; e < ma * nb; e += blockDim.x) {
                    int r = e % ma, c = e / ma;
                    X[relK[r] + c * LDS] += MvK[r + (int64_t)(cb + c) * ma];
                }
                __syncthreads();
            }
            for (int e = tid; e < nn * nb; e += blockDim.x) {
                int r = e % nn, c = e / nn;
                X[r + c * LDS] += S.B[f.c0 + r + (int64_t)(cb + c) * S.ldb];
            }
            __syncthreads();
            for (int k = 0; k < nn; ++k) {
                if (!unit && tid < nb) X[k + tid * LDS] /= Ps[k + k * LDS];
                __syncthreads();
                const int m = nj - k - 1;
                for (int e = tid; e < m * nb; e += blockDim.x) {
                    int r = k + 1 + e % m, c = e / m;
                    X[r + c * LDS] -= Ps[r + k * LDS] * X[k + c * LDS];
                }
                __syncthreads();
            }
            for (int e = tid; e < nj * nb; e += blockDim.x) {
                int r = e % nj, c = e / nj;
                if (r < nn) S.B[f.c0 + r + (int64_t)(cb + c) * S.ldb] = X[r + c * LDS];
                else MvJ[(r - nn) + (int64_t)(cb + c) * na] = X[r + c * LDS];
            }
            __syncthreads();
        }
        release_agent();
        __syncthreads();
        if (!persistent) return;
        if (tid == 0) __hip_atomic_store(F.flags + J, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_ssmall_bwd(DevPlan P, DevFactor<T> F, DevSolve<T> S,
                                                    const int32_t* list, int count,
                                                    int persistent) {
    __shared__ T Ps[SMALL_NJ * LDS];
    __shared__ T X[LDS * RB];
    __shared__ int s_t;
    const int tid = threadIdx.x;
    const int unit = F.diag != 0;
    int t = blockIdx.x;
    for (;;) {
        if (persistent) {
            if (tid == 0) s_t = (int)atomicAdd(F.ticket, 1u);
            __syncthreads();
            t = s_t;
        }
        if (t >= count) return;
        const int J = persistent ? list[count - 1 - t] : list[t];
        const Front f = P.fronts[J];
        const int nn = f.nn, na = f.na, nj = nn + na;
        if (tid == 0 && persistent && f.parent >= 0 && P.in_small[f.parent])
            wait_flag(F.flags + f.parent, F.hang);
        __syncthreads();
        acquire_agent();
        const T* Pj = F.panel + f.poff;
        for (int e = tid; e < nj * nn; e += blockDim.x) {
            int r = e % nj, c = e / nj;
            Ps[r + c * LDS] = Pj[e];
        }
        const int32_t* rows = P.rows + f.rowoff;
        for (int cb = 0; cb < S.nrhs; cb += RB) {
            const int nb = min(RB, S.nrhs - cb);
            for (int e = tid; e < nj * nb; e += blockDim.x) {
                int r = e % nj, c = e / nj;
                X[r + c * LDS] = S.B[rows[r] + (int64_t)(cb + c) * S.ldb];
            }
            __syncthreads();
            for (int e = tid; e < nn * nb; e += blockDim.x) {
                int k = e % nn, c = e / nn;
                T s = (T)0;
                for (int r = nn; r < nj; ++r) s += Ps[r + k * LDS] * X[r + c * LDS];
                X[k + c * LDS] -= s;
            }
            __syncthreads();
            for (int k = nn - 1; k >= 0; --k) {
                if (!unit && tid < nb) X[k + tid * LDS] /= Ps[k + k * LDS];
                __syncthreads();
                for (int e = tid; e < k * nb; e += blockDim.x) {
                    int j = e % k, c = e / k;
                    X[j + c * LDS] -= Ps[k + j * LDS] * X[k + c * LDS];
                }
                __syncthreads();
            }
            for (int e = tid; e < nn * nb; e += blockDim.x) {
                int r = e % nn, c = e / nn;
                S.B[f.c0 + r + (int64_t)(cb + c) * S.ldb] = X[r + c * LDS];
            }
            __syncthreads();
        }
        release_agent();
        __syncthreads();
        if (!persistent) return;
        if (tid == 0) __hip_atomic_store(F.flags + J, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Forward gather for big fronts: rows [r0, r1) of front J receive the
// children's update vectors; rows < nn are added to C, rows >= nn
// initialize M_J.
template <typename T>
__global__ __launch_bounds__(256) void k_sgather(DevPlan P, DevSolve<T> S, int t0) {
    __shared__ T acc[EA_CHUNK * RB];
    const SolveTask st = P.stask[t0 + blockIdx.x];
    const Front f = P.fronts[st.front];
    const int r0 = st.r0, r1 = st.r1, len = r1 - r0, tid = threadIdx.x;
    T* MvJ = S.Mv + P.mvoff1[st.front] * S.nrhs;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        for (int e = tid; e < EA_CHUNK * RB; e += blockDim.x) acc[e] = (T)0;
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
            const int cnt = lo - a;
            const T* MvK = S.Mv + P.mvoff1[K] * S.nrhs;
            for (int e = tid; e < cnt * nb; e += blockDim.x) {
                int i = a + e % cnt, c = e / cnt;
                acc[(relK[i] - r0) + c * EA_CHUNK] += MvK[i + (int64_t)(cb + c) * ma];
            }
            __syncthreads();
        }
        for (int e = tid; e < len * nb; e += blockDim.x) {
            int r = r0 + e % len, c = e / len;
            T v = acc[(r - r0) + c * EA_CHUNK];
            if (r < f.nn) S.B[f.c0 + r + (int64_t)(cb + c) * S.ldb] += v;
            else MvJ[(r - f.nn) + (int64_t)(cb + c) * f.na] = v;
        }
        __syncthreads();
    }
}

// X_k <- W_k X_k (forward) or W_k' X_k (backward), X_k = C rows of panel k.
template <typename T>
__global__ __launch_bounds__(256) void k_sdiag(DevPlan P, DevFactor<T> F, DevSolve<T> S, int t0,
                                               int backward) {
    __shared__ T Ws[NB * (NB + 1)];
    __shared__ T X[NB * RB];
    const SolveTask st = P.stask[t0 + blockIdx.x];
    const Front f = P.fronts[st.front];
    const int kb = st.kb, tid = threadIdx.x;
    const T* W = F.W + f.woff + st.k0;
    for (int e = tid; e < kb * kb; e += blockDim.x) {
        int r = e % kb, c = e / kb;
        Ws[r + c * (NB + 1)] = W[r + (int64_t)c * f.nn];
    }
    T* Xg = S.B + f.c0 + st.k0;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        for (int e = tid; e < kb * nb; e += blockDim.x) {
            int r = e % kb, c = e / kb;
            X[r + c * NB] = Xg[r + (int64_t)(cb + c) * S.ldb];
        }
        __syncthreads();
        for (int e = tid; e < kb * nb; e += blockDim.x) {
            int r = e % kb, c = e / kb;
            T s = (T)0;
            if (!backward) {
                for (int k = 0; k <= r; ++k) s += Ws[r + k * (NB + 1)] * X[k + c * NB];
            } else {
                for (int k = r; k < kb; ++k) s += Ws[k + r * (NB + 1)] * X[k + c * NB];
            }
            Xg[r + (int64_t)(cb + c) * S.ldb] = s;
        }
        __syncthreads();
    }
}

// Forward update: rows [r0, r1) of front J: y -= P[r, k0:k0+kb] X_k.
template <typename T>
__global__ __launch_bounds__(256) void k_supd_fwd(DevPlan P, DevFactor<T> F, DevSolve<T> S,
                                                  int t0) {
    __shared__ T X[NB * RB];
    const SolveTask st = P.stask[t0 + blockIdx.x];
    const Front f = P.fronts[st.front];
    const int nj = f.nn + f.na, kb = st.kb, tid = threadIdx.x;
    const int r = st.r0 + (tid & 63), cg = tid >> 6;
    const T* Pk = F.panel + f.poff + (int64_t)st.k0 * nj;
    const T* Xg = S.B + f.c0 + st.k0;
    T* MvJ = S.Mv + P.mvoff1[st.front] * S.nrhs;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        for (int e = tid; e < kb * nb; e += blockDim.x) {
            int k = e % kb, c = e / kb;
            X[k + c * NB] = Xg[k + (int64_t)(cb + c) * S.ldb];
        }
        __syncthreads();
        if (r < st.r1) {
            T acc[RB / 4];
#pragma unroll
            for (int i = 0; i < RB / 4; ++i) acc[i] = (T)0;
            for (int k = 0; k < kb; ++k) {
                const T p = Pk[r + (int64_t)k * nj];
#pragma unroll
                for (int i = 0; i < RB / 4; ++i) acc[i] += p * X[k + (cg + 4 * i) * NB];
            }
#pragma unroll
            for (int i = 0; i < RB / 4; ++i) {
                const int c = cg + 4 * i;
                if (c < nb) {
                    if (r < f.nn) S.B[f.c0 + r + (int64_t)(cb + c) * S.ldb] -= acc[i];
                    else MvJ[(r - f.nn) + (int64_t)(cb + c) * f.na] -= acc[i];
                }
            }
        }
        __syncthreads();
    }
}

// Backward gather: local columns [r0, r1) of front J:
//   x_c -= sum_{r >= nn} P[r, c] C[rows[r]].
template <typename T>
__global__ __launch_bounds__(256) void k_sbgather(DevPlan P, DevFactor<T> F, DevSolve<T> S,
                                                  int t0) {
    __shared__ T Pt[64 * 65];
    __shared__ T Xs[64 * RB];
    const SolveTask st = P.stask[t0 + blockIdx.x];
    const Front f = P.fronts[st.front];
    const int nj = f.nn + f.na, tid = threadIdx.x;
    const int ncol = st.r1 - st.r0;
    const int32_t* rows = P.rows + f.rowoff;
    const T* Pj = F.panel + f.poff;
    const int cl = tid & 63, cg = tid >> 6;
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        T acc[RB / 4];
#pragma unroll
        for (int i = 0; i < RB / 4; ++i) acc[i] = (T)0;
        for (int rt = f.nn; rt < nj; rt += 64) {
            const int nr = min(64, nj - rt);
            for (int e = tid; e < 64 * ncol; e += blockDim.x) {
                int r = e & 63, c = e >> 6;
                Pt[r + c * 65] = r < nr ? Pj[(rt + r) + (int64_t)(st.r0 + c) * nj] : (T)0;
            }
            for (int e = tid; e < 64 * nb; e += blockDim.x) {
                int r = e & 63, c = e >> 6;
                Xs[r + c * 64] = r < nr ? S.B[rows[rt + r] + (int64_t)(cb + c) * S.ldb] : (T)0;
            }
            __syncthreads();
            if (cl < ncol) {
                for (int r = 0; r < nr; ++r) {
                    const T p = Pt[r + cl * 65];
#pragma unroll
                    for (int i = 0; i < RB / 4; ++i) acc[i] += p * Xs[r + (cg + 4 * i) * 64];
                }
            }
            __syncthreads();
        }
        if (cl < ncol) {
#pragma unroll
            for (int i = 0; i < RB / 4; ++i) {
                const int c = cg + 4 * i;
                if (c < nb) S.B[f.c0 + st.r0 + cl + (int64_t)(cb + c) * S.ldb] -= acc[i];
            }
        }
        __syncthreads();
    }
}

// Backward update: local columns [r0, r1) (< k0): x_j -= P[k0:k0+kb, j]' X_k.
template <typename T>
__global__ __launch_bounds__(256) void k_supd_bwd(DevPlan P, DevFactor<T> F, DevSolve<T> S,
                                                  int t0) {
    __shared__ T Pt[NB * 65];
    __shared__ T X[NB * RB];
    const SolveTask st = P.stask[t0 + blockIdx.x];
    const Front f = P.fronts[st.front];
    const int nj = f.nn + f.na, kb = st.kb, tid = threadIdx.x;
    const int ncol = st.r1 - st.r0;
    const T* Pj = F.panel + f.poff;
    const T* Xg = S.B + f.c0 + st.k0;
    const int cl = tid & 63, cg = tid >> 6;
    for (int e = tid; e < NB * ncol; e += blockDim.x) {
        int k = e & 63, c = e >> 6;
        Pt[k + c * 65] = k < kb ? Pj[(st.k0 + k) + (int64_t)(st.r0 + c) * nj] : (T)0;
    }
    for (int cb = 0; cb < S.nrhs; cb += RB) {
        const int nb = min(RB, S.nrhs - cb);
        for (int e = tid; e < kb * nb; e += blockDim.x) {
            int k = e % kb, c = e / kb;
            X[k + c * NB] = Xg[k + (int64_t)(cb + c) * S.ldb];
        }
        __syncthreads();
        if (cl < ncol) {
            T acc[RB / 4];
#pragma unroll
            for (int i = 0; i < RB / 4; ++i) acc[i] = (T)0;
            for (int k = 0; k < kb; ++k) {
                const T p = Pt[k + cl * 65];
#pragma unroll
                for (int i = 0; i < RB / 4; ++i) acc[i] += p * X[k + (cg + 4 * i) * NB];
            }
#pragma unroll
            for (int i = 0; i < RB / 4; ++i) {
                const int c = cg + 4 * i;
                if (c < nb) S.B[f.c0 + st.r0 + cl + (int64_t)(cb + c) * S.ldb] -= acc[i];
            }
        }
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

}  // namespace

template <typename T>
void launch_solve_small_fwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                            const int32_t* list, int count, bool persistent, hipStream_t s) {
    if (count <= 0) return;
    int grid = persistent ? (count < 2048 ? count : 2048) : count;
    hipLaunchKernelGGL(k_ssmall_fwd<T>, dim3(grid), dim3(256), 0, s, P, F, S, list, count,
                       persistent ? 1 : 0);
}
template <typename T>
void launch_solve_small_bwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S,
                            const int32_t* list, int count, bool persistent, hipStream_t s) {
    if (count <= 0) return;
    int grid = persistent ? (count < 2048 ? count : 2048) : count;
    hipLaunchKernelGGL(k_ssmall_bwd<T>, dim3(grid), dim3(256), 0, s, P, F, S, list, count,
                       persistent ? 1 : 0);
}
template <typename T>
void launch_sgather(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, int t0,
                    int t1, hipStream_t s) {
    (void)F;
    if (t1 <= t0) return;
    hipLaunchKernelGGL(k_sgather<T>, dim3(t1 - t0), dim3(256), 0, s, P, S, t0);
}
template <typename T>
void launch_sdiag(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, int t0,
                  int t1, bool backward, hipStream_t s) {
    if (t1 <= t0) return;
    hipLaunchKernelGGL(k_sdiag<T>, dim3(t1 - t0), dim3(256), 0, s, P, F, S, t0,
                       backward ? 1 : 0);
}
template <typename T>
void launch_supd_fwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, int t0,
                     int t1, hipStream_t s) {
    if (t1 <= t0) return;
    hipLaunchKernelGGL(k_supd_fwd<T>, dim3(t1 - t0), dim3(256), 0, s, P, F, S, t0);
}
template <typename T>
void launch_sbgather(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, int t0,
                     int t1, hipStream_t s) {
    if (t1 <= t0) return;
    hipLaunchKernelGGL(k_sbgather<T>, dim3(t1 - t0), dim3(256), 0, s, P, F, S, t0);
}
template <typename T>
void launch_supd_bwd(const DevPlan& P, const DevFactor<T>& F, const DevSolve<T>& S, int t0,
                     int t1, hipStream_t s) {
    if (t1 <= t0) return;
    hipLaunchKernelGGL(k_supd_bwd<T>, dim3(t1 - t0), dim3(256), 0, s, P, F, S, t0);
}
template <typename T>
void launch_dscale(const DevFactor<T>& F, const DevSolve<T>& S, int n, hipStream_t s) {
    int64_t tot = (int64_t)n * S.nrhs;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_dscale<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       (const T*)F.d, S.B, S.ldb, n, S.nrhs);
}
template <typename TB>
void launch_permute(const double* src, int64_t lds, TB* dst, int64_t ldd, const int64_t* perm,
                    int n, int nrhs, bool inverse, hipStream_t s) {
    int64_t tot = (int64_t)n * nrhs;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_permute<TB>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, src,
                       lds, dst, ldd, perm, n, nrhs, inverse ? 1 : 0);
}
void launch_permute_back(const void* src, int64_t lds, bool src_f32, double* dst, int64_t ldd,
                         const int64_t* perm, int n, int nrhs, hipStream_t s) {
    int64_t tot = (int64_t)n * nrhs;
    if (tot <= 0) return;
    if (src_f32)
        hipLaunchKernelGGL(k_permute_back<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256),
                           0, s, (const float*)src, lds, dst, ldd, perm, n, nrhs);
    else
        hipLaunchKernelGGL(k_permute_back<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256),
                           0, s, (const double*)src, lds, dst, ldd, perm, n, nrhs);
}

#define MF_SINST(T)                                                                              \
    template void launch_solve_small_fwd<T>(const DevPlan&, const DevFactor<T>&,                 \
                                            const DevSolve<T>&, const int32_t*, int, bool,       \
                                            hipStream_t);                                        \
    template void launch_solve_small_bwd<T>(const DevPlan&, const DevFactor<T>&,                 \
                                            const DevSolve<T>&, const int32_t*, int, bool,       \
                                            hipStream_t);                                        \
    template void launch_sgather<T>(const DevPlan&, const DevFactor<T>&, const DevSolve<T>&, int, \
                                    int, hipStream_t);                                           \
    template void launch_sdiag<T>(const DevPlan&, const DevFactor<T>&, const DevSolve<T>&, int,  \
                                  int, bool, hipStream_t);                                       \
    template void launch_supd_fwd<T>(const DevPlan&, const DevFactor<T>&, const DevSolve<T>&,    \
                                     int, int, hipStream_t);                                     \
    template void launch_sbgather<T>(const DevPlan&, const DevFactor<T>&, const DevSolve<T>&,    \
                                     int, int, hipStream_t);                                     \
    template void launch_supd_bwd<T>(const DevPlan&, const DevFactor<T>&, const DevSolve<T>&,    \
                                     int, int, hipStream_t);                                     \
    template void launch_dscale<T>(const DevFactor<T>&, const DevSolve<T>&, int, hipStream_t);   \
    template void launch_permute<T>(const double*, int64_t, T*, int64_t, const int64_t*, int,    \
                                    int, bool, hipStream_t);
MF_SINST(double)
MF_SINST(float)

}  // namespace mf
