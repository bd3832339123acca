// mf_factor.hip — gfx950 kernels of the multifrontal numeric factorization.
//
//   k_small     fused front kernel for internal fronts with nJ <= 64: the
//               front lives in LDS; extend-add of the children's update
//               matrices, assembly, partial factorization and the write-out
//               of the panel and of U in one workgroup.  Persistent mode
//               hands out fronts by ticket in level order and waits on the
//               children's completion flags (bounded wait).
//               Reference: chol_loop_snd!/nod! cholesky.jl:434-692,
//               chol_send! cholesky.jl:950-983, potrf2! blas/potrf.jl:3-41.
//   k_ea        extend-add of big fronts, one task per parent column: the
//               column is accumulated in LDS over all children (in the
//               reference's reverse child order) and then added to the panel
//               (F11/F21 += F, cholesky.jl:520-521) or written to U (M22 <-
//               F22, cholesky.jl:542).
//   k_diag      factor + invert one 64x64 diagonal block (potrf2!).
//   k_gemm      batched 64x64-tile MFMA GEMM: panel solve (trsm!,
//               blas/trsx.jl:156-159, as a product with the inverted
//               diagonal block), trailing updates and the update matrix
//               M22 -= L21 [D] L21' (syrk!, blas/syrk.jl:137-159).
//   k_copyout   internal panels -> the reference's Dval/Lval/d layout.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mf_kernels.h"

namespace mf {

namespace {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // D register v of lane l holds row (l/16) + 4v, column l%16
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // D register v of lane l holds row 4(l/16) + v, column l%16
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

template <typename T>
__device__ __forceinline__ double sign_of(const DevFactor<T>& F, int gcol) {
    if (F.signs) return F.signs[gcol];
    return F.diag == 0 ? 1.0 : 0.0;
}

// regularize(R, S, Djj, j): no_regularization.jl:27-33 and
// dynamic_regularization.jl:117-123.  0 signals a failed pivot.
template <typename T>
__device__ __forceinline__ double regularize(const DevFactor<T>& F, double s, double Djj) {
    if (F.reg_kind == 1) return (s * Djj < F.eps) ? F.delta * s : Djj;
    if (s != 0.0 && !(s * Djj > 0.0)) return 0.0;
    return Djj;
}

__device__ __forceinline__ void release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Bounded wait for a completion flag; gives up (and records it) after ~1 s.
__device__ __forceinline__ void wait_flag(int32_t* flag, int32_t* hang) {
    int spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
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

constexpr int LDS = SMALL_NJ + 1;  // leading dimension of LDS fronts

// Right-looking partial factorization of the first nn columns of the
// nj x nj front M (lower triangle, leading dimension LDS) in LDS.
// Returns after a failed pivot (recorded in F.info).
template <typename T>
__device__ void factor_lds(T* M, int nn, int nj, int gc0, const DevFactor<T>& F,
                           T* s_piv, T* s_inv, int* s_fail) {
    const int tid = threadIdx.x;
    for (int k = 0; k < nn; ++k) {
        if (tid == 0) {
            double akk = (double)M[k + k * LDS];
            double piv = regularize(F, sign_of(F, gc0 + k), akk);
            if (piv == 0.0) {
                *s_fail = 1;
                atomicMin(F.info, (unsigned long long)(gc0 + k + 1));
            } else if (F.diag == 0) {
                T r = (T)sqrt(piv);
                M[k + k * LDS] = r;
                *s_piv = (T)1;
                *s_inv = (T)1 / r;
            } else {
                F.d[gc0 + k] = (T)piv;
                *s_piv = (T)piv;
                *s_inv = (T)1 / (T)piv;
            }
        }
        __syncthreads();
        if (*s_fail) return;
        const T inv = *s_inv;
        const int r = tid & 63, cg = tid >> 6;
        if (tid < 64 && r > k && r < nj) M[r + k * LDS] *= inv;
        __syncthreads();
        const T piv = *s_piv;
        if (r < nj) {
            const T lr = M[r + k * LDS] * piv;
            for (int c = k + 1 + cg; c <= r; c += 4) M[r + c * LDS] -= lr * M[c + k * LDS];
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_small(DevPlan P, DevFactor<T> F, const int32_t* list,
                                               int count, int persistent) {
    __shared__ T M[SMALL_NJ * LDS];
    __shared__ T s_piv, s_inv;
    __shared__ int s_fail, s_t;
    const int tid = threadIdx.x;
    int t = blockIdx.x;
    for (;;) {
        if (persistent) {
            if (tid == 0) s_t = (int)atomicAdd(F.ticket, 1u);
            __syncthreads();
            t = s_t;
            __syncthreads();
        }
        if (t >= count) break;
        const int J = list[t];
        const Front f = P.fronts[J];
        const int nn = f.nn, na = f.na, nj = nn + na;
        if (tid == 0) {
            s_fail = 0;
            if (persistent)
                for (int q = f.chd0; q < f.chd1; ++q) wait_flag(F.flags + P.chd[q], F.hang);
        }
        __syncthreads();
        acquire_agent();
        const int r = tid & 63, cg = tid >> 6;
        for (int c = cg; c < nj; c += 4) M[r + c * LDS] = (T)0;
        __syncthreads();
        for (int q = f.chd0; q < f.chd1; ++q) {
            const Front kf = P.fronts[P.chd[q]];
            const int ma = kf.na;
            const int32_t* relK = P.rel + kf.reloff;
            const T* UK = F.U + kf.uoff;
            if (r < ma) {
                const int rr = relK[r];
                for (int c = cg; c <= r; c += 4) M[rr + relK[c] * LDS] += UK[r + c * ma];
            }
            __syncthreads();
        }
        T* Pj = F.panel + f.poff;
        if (r < nj)
            for (int c = cg; c <= r && c < nn; c += 4) M[r + c * LDS] += Pj[r + c * nj];
        __syncthreads();
        factor_lds<T>(M, nn, nj, f.c0, F, &s_piv, &s_inv, &s_fail);
        if (r < nj)
            for (int c = cg; c <= r && c < nn; c += 4) Pj[r + c * nj] = M[r + c * LDS];
        if (na > 0 && r < na) {
            T* Uj = F.U + f.uoff;
            for (int c = cg; c <= r; c += 4) Uj[r + c * na] = M[(nn + r) + (nn + c) * LDS];
        }
        release_agent();
        __syncthreads();
        if (!persistent) break;
        if (tid == 0) __hip_atomic_store(F.flags + J, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        // closes the thread-0 block before the back-edge: every thread of the
        // workgroup reaches the next ticket together
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_ea(DevPlan P, DevFactor<T> F, int t0) {
    __shared__ T acc[EA_CHUNK];
    const EaTask e = P.ea[t0 + blockIdx.x];
    const Front f = P.fronts[e.front];
    const int nj = f.nn + f.na, pc = e.col, tid = threadIdx.x;
    T* Pj = F.panel + f.poff;
    for (int r0 = e.row0; r0 < e.row1; r0 += EA_CHUNK) {
        const int r1 = min(r0 + EA_CHUNK, e.row1);
        acc[tid] = (T)0;
        __syncthreads();
        for (int s = e.src0; s < e.src1; ++s) {
            const Front kf = P.fronts[P.ea_src[2 * s]];
            const int tcol = P.ea_src[2 * s + 1];
            const int ma = kf.na;
            const int32_t* relK = P.rel + kf.reloff;
            // rows tcol..ma-1 of the child column land on parent rows relK[.] (increasing)
            int lo = tcol, hi = ma;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (relK[mid] < r0) lo = mid + 1; else hi = mid;
            }
            int a = lo;
            hi = ma;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (relK[mid] < r1) lo = mid + 1; else hi = mid;
            }
            const T* UK = F.U + kf.uoff + (int64_t)tcol * ma;
            for (int i = a + tid; i < lo; i += blockDim.x) acc[relK[i] - r0] += UK[i];
            __syncthreads();
        }
        const int r = r0 + tid;
        if (r < r1) {
            if (pc < f.nn) Pj[r + (int64_t)pc * nj] += acc[tid];
            else F.U[f.uoff + (r - f.nn) + (int64_t)(pc - f.nn) * f.na] = acc[tid];
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double rdlane(double x, int l) {
    long long b = __double_as_longlong(x);
    int lo = __builtin_amdgcn_readlane((int)b, l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float rdlane(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// One wavefront factors and inverts a kb x kb (kb <= 64) diagonal block.
// Lane i keeps row i of the block in registers; the scaled column k is
// broadcast through LDS, so every step is 63-k independent FMAs per lane.
// The factorization is potrf2!'s right-looking column loop
// (blas/potrf.jl:3-41); the inverse W = L^-1 (unit L for LDLt) is formed by
// right-looking forward substitution with lane c owning column c, and is
// stored (ld = nn) for the panel solve and the triangular solves.
template <typename T>
__global__ __launch_bounds__(64) void k_diag(DevPlan P, DevFactor<T> F, int t0, int invert_only) {
    __shared__ T Tr[NB * (NB + 1)];
    __shared__ T colk[NB];
    const DiagTask dt = P.diag[t0 + blockIdx.x];
    const Front f = P.fronts[dt.front];
    const int nj = f.nn + f.na, kb = dt.kb, k0 = dt.k0, lane = threadIdx.x;
    T* blk = F.panel + f.poff + k0 + (int64_t)k0 * nj;
    T a[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        a[c] = (lane < kb && c < kb && c <= lane) ? blk[lane + (int64_t)c * nj] : (T)0;
    if (!invert_only) {
        bool failed = false;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (k < kb) {
                const int g = f.c0 + k0 + k;
                const T akk = rdlane(a[k], k);
                double piv = regularize(F, sign_of(F, g), (double)akk);
                if (!failed && piv == 0.0) {
                    failed = true;
                    if (lane == 0) atomicMin(F.info, (unsigned long long)(g + 1));
                }
                if (piv == 0.0) piv = 1.0;  // keeps the rest finite after a failure
                T dk, fm;
                if (F.diag == 0) {
                    dk = (T)sqrt(piv);
                    fm = (T)1;
                    if (lane == k) a[k] = dk;
                } else {
                    dk = (T)piv;
                    fm = dk;
                    if (lane == 0 && !failed) F.d[g] = dk;
                }
                const T inv = (T)1 / dk;
                const T lik = (lane > k) ? a[k] * inv : (T)0;
                if (lane > k) a[k] = lik;
                colk[lane] = lik;
                __syncthreads();
                const T lf = lik * fm;
#pragma unroll
                for (int j = k + 1; j < NB; ++j) {
                    const T ljk = colk[j];
                    if (lane >= j) a[j] -= lf * ljk;
                }
                __syncthreads();
            }
        }
        if (lane < kb) {
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c < kb && c <= lane) blk[lane + (int64_t)c * nj] = a[c];
        }
    }
    // W = L^-1, lane c owns column c: w[i] = W[i][c]
    const bool unit = F.diag != 0;
    T w[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) w[i] = (lane == i) ? (T)1 : (T)0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (k < kb) {
            if (!unit) w[k] *= (T)1 / rdlane(a[k], k);
            colk[lane] = a[k];  // column k of L (lane i holds L[i][k])
            __syncthreads();
            const T wk = w[k];
#pragma unroll
            for (int i = k + 1; i < NB; ++i) w[i] -= colk[i] * wk;
            __syncthreads();
        }
    }
    // transpose through LDS so that the column-major stores are coalesced
#pragma unroll
    for (int i = 0; i < NB; ++i) Tr[i * (NB + 1) + lane] = (i < kb) ? w[i] : (T)0;
    __syncthreads();
    T* W = F.W + f.woff + k0;
    if (lane < kb) {
        for (int c = 0; c < kb; ++c) W[lane + (int64_t)c * f.nn] = Tr[lane * (NB + 1) + c];
    }
}

template <typename T>
__device__ __forceinline__ T* arr_base(const DevFactor<T>& F, int a) {
    return a == ARR_PANEL ? F.panel : (a == ARR_U ? F.U : F.W);
}

// 64x64 output tile per workgroup; wave w owns a 32x32 quarter as 2x2 blocks
// of 16x16 MFMA tiles (v_mfma_*_16x16x4: lane l holds A[l%16][l/16] and
// B[l/16][l%16]; the row of output register v is Mfma<T>::row).
template <typename T>
__global__ __launch_bounds__(256) void k_gemm(DevFactor<T> F, const GemmTask* tasks, int t0) {
    typedef typename Mfma<T>::v4 v4;
    constexpr int KC = 16;
    constexpr int LDT = 64 + 4;
    __shared__ T As[KC * LDT];
    __shared__ T Bs[KC * LDT];
    const GemmTask g = tasks[t0 + blockIdx.x];
    const T* A = arr_base(F, g.arrA) + g.a_off;
    const T* B = arr_base(F, g.arrB) + g.b_off;
    T* C = arr_base(F, g.arrC) + g.c_off;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    const T* s = (F.diag != 0 && g.mode == G_SUB && g.s_col >= 0) ? F.d + g.s_col : nullptr;
    v4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = v4{0, 0, 0, 0};
    for (int k0 = 0; k0 < g.k; k0 += KC) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = tid + 256 * i, r = e & 63, kk = e >> 6, gk = k0 + kk;
            T av = (T)0, bv = (T)0;
            if (gk < g.k) {
                if (r < g.m) {
                    av = A[r + (int64_t)gk * g.lda];
                    if (s) av *= s[gk];
                }
                if (r < g.n) bv = B[r + (int64_t)gk * g.ldb];
            }
            As[kk * LDT + r] = av;
            Bs[kk * LDT + r] = bv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            const T a0 = As[(kk + lk) * LDT + wr + li];
            const T a1 = As[(kk + lk) * LDT + wr + 16 + li];
            const T b0 = Bs[(kk + lk) * LDT + wc + li];
            const T b1 = Bs[(kk + lk) * LDT + wc + 16 + li];
            acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
            acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
            acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
            acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
    const bool scale = (g.mode == G_TRSM) && F.diag != 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wr + i * 16 + Mfma<T>::row(lk, v);
                const int col = wc + j * 16 + li;
                if (row < g.m && col < g.n && (!g.lower || row >= col)) {
                    T* cp = C + row + (int64_t)col * g.ldc;
                    if (g.mode == G_SUB) {
                        *cp -= acc[i][j][v];
                    } else {
                        T x = acc[i][j][v];
                        if (scale) x /= F.d[g.s_col + col];
                        *cp = x;
                    }
                }
            }
        }
    }
}

template <typename T>
__global__ void k_fill(T* p, int64_t count, T value) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = value;
}

// copyto!-layout Dval/Lval (assembled A) -> internal panels (zeroed before);
// one workgroup per global column.
template <typename T>
__global__ __launch_bounds__(256) void k_assemble_ref(DevPlan P, DevFactor<T> F,
                                                      const double* Dval, const double* Lval) {
    const int g = blockIdx.x;
    const OutFront o = P.out[P.col2ref[g]];
    const Front f = P.fronts[o.J];
    const int nj = f.nn + f.na, c = g - o.r0, cl = o.r0 - f.c0;
    T* col = F.panel + f.poff + (int64_t)(cl + c) * nj;
    for (int r = c + threadIdx.x; r < o.nn; r += blockDim.x)
        col[cl + r] = (T)Dval[o.Dp + r + (int64_t)c * o.nn];
    for (int i = threadIdx.x; i < o.na; i += blockDim.x)
        col[P.septpos[o.sepoff + i]] = (T)Lval[o.Lp + i + (int64_t)c * o.na];
}

// setflatindex! for every nonzero: panel slot and reference slot.
template <typename T>
__global__ void k_refresh(DevFactor<T> F, const int64_t* tp, const int64_t* tr, const double* x,
                          int64_t nnz, double* Dref, int64_t ndz, double* Lref) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const double v = x[k];
    const int64_t p = tp[k];
    if (p >= 0) F.panel[p] = (T)v;
    const int64_t q = tr[k];
    if (q >= 0) {
        if (q < ndz) Dref[q] = v;
        else Lref[q - ndz] = v;
    }
}

// Internal panels -> the reference's Dval/Lval/d; one workgroup per global column.
template <typename T>
__global__ __launch_bounds__(256) void k_copyout(DevPlan P, DevFactor<T> F, double* Dval,
                                                 double* Lval, double* dout) {
    const int g = blockIdx.x;
    const OutFront o = P.out[P.col2ref[g]];
    const Front f = P.fronts[o.J];
    const int nj = f.nn + f.na, c = g - o.r0, cl = o.r0 - f.c0;
    const T* col = F.panel + f.poff + (int64_t)(cl + c) * nj;
    const bool keep_diag = (F.diag != 0 && o.nn == 1);  // chol_loop_nod!: Dval keeps A's entry
    if (Dval && !keep_diag) {
        for (int r = threadIdx.x; r < o.nn; r += blockDim.x)
            Dval[o.Dp + r + (int64_t)c * o.nn] = r >= c ? (double)col[cl + r] : 0.0;
    }
    if (Lval) {
        for (int i = threadIdx.x; i < o.na; i += blockDim.x)
            Lval[o.Lp + i + (int64_t)c * o.na] = (double)col[P.septpos[o.sepoff + i]];
    }
    if (dout && F.diag != 0 && threadIdx.x == 0) dout[g] = (double)F.d[g];
}

inline int grid_for(int64_t count, int block) {
    int64_t g = (count + block - 1) / block;
    if (g > 65536) g = 65536;
    return (int)(g < 1 ? 1 : g);
}

}  // namespace

template <typename T>
void launch_fill(T* p, int64_t count, T value, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(count, 256)), dim3(256), 0, s, p, count, value);
}

template <typename T>
void launch_assemble_ref(const DevPlan& P, const DevFactor<T>& F, const double* Dval,
                         const double* Lval, int nfr, hipStream_t s) {
    if (nfr <= 0 || P.n <= 0) return;
    hipLaunchKernelGGL(k_assemble_ref<T>, dim3(P.n), dim3(256), 0, s, P, F, Dval, Lval);
}

template <typename T>
void launch_refresh(const DevFactor<T>& F, const int64_t* tp, const int64_t* tr,
                    const double* nzval, int64_t nnz, double* Dref, int64_t ndz, double* Lref,
                    hipStream_t s) {
    if (nnz <= 0) return;
    hipLaunchKernelGGL(k_refresh<T>, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s, F, tp,
                       tr, nzval, nnz, Dref, ndz, Lref);
}

template <typename T>
void launch_small(const DevPlan& P, const DevFactor<T>& F, const int32_t* list, int count,
                  bool persistent, hipStream_t s) {
    if (count <= 0) return;
    int grid = persistent ? (count < 2048 ? count : 2048) : count;
    hipLaunchKernelGGL(k_small<T>, dim3(grid), dim3(256), 0, s, P, F, list, count,
                       persistent ? 1 : 0);
}

template <typename T>
void launch_ea(const DevPlan& P, const DevFactor<T>& F, int t0, int t1, hipStream_t s) {
    if (t1 <= t0) return;
    hipLaunchKernelGGL(k_ea<T>, dim3(t1 - t0), dim3(256), 0, s, P, F, t0);
}

template <typename T>
void launch_diag(const DevPlan& P, const DevFactor<T>& F, int t0, int t1, bool invert_only,
                 hipStream_t s) {
    if (t1 <= t0) return;
    hipLaunchKernelGGL(k_diag<T>, dim3(t1 - t0), dim3(64), 0, s, P, F, t0, invert_only ? 1 : 0);
}

template <typename T>
void launch_gemm(const DevPlan& P, const DevFactor<T>& F, int t0, int t1, hipStream_t s) {
    if (t1 <= t0) return;
    hipLaunchKernelGGL(k_gemm<T>, dim3(t1 - t0), dim3(256), 0, s, F, P.gemm, t0);
}

template <typename T>
void launch_copyout(const DevPlan& P, const DevFactor<T>& F, double* Dval, double* Lval,
                    double* dout, int nfr, int n, hipStream_t s) {
    if (nfr <= 0 || n <= 0) return;
    hipLaunchKernelGGL(k_copyout<T>, dim3(n), dim3(256), 0, s, P, F, Dval, Lval, dout);
}

#define MF_INST(T)                                                                              \
    template void launch_fill<T>(T*, int64_t, T, hipStream_t);                                  \
    template void launch_assemble_ref<T>(const DevPlan&, const DevFactor<T>&, const double*,    \
                                         const double*, int, hipStream_t);                      \
    template void launch_refresh<T>(const DevFactor<T>&, const int64_t*, const int64_t*,        \
                                    const double*, int64_t, double*, int64_t, double*,          \
                                    hipStream_t);                                               \
    template void launch_small<T>(const DevPlan&, const DevFactor<T>&, const int32_t*, int,     \
                                  bool, hipStream_t);                                           \
    template void launch_ea<T>(const DevPlan&, const DevFactor<T>&, int, int, hipStream_t);     \
    template void launch_diag<T>(const DevPlan&, const DevFactor<T>&, int, int, bool,           \
                                 hipStream_t);                                                 \
    template void launch_gemm<T>(const DevPlan&, const DevFactor<T>&, int, int, hipStream_t);   \
    template void launch_copyout<T>(const DevPlan&, const DevFactor<T>&, double*, double*,      \
                                    double*, int, int, hipStream_t);
MF_INST(double)
MF_INST(float)
template void launch_fill<int32_t>(int32_t*, int64_t, int32_t, hipStream_t);
template void launch_fill<uint32_t>(uint32_t*, int64_t, uint32_t, hipStream_t);
template void launch_fill<unsigned long long>(unsigned long long*, int64_t, unsigned long long,
                                              hipStream_t);

}  // namespace mf
