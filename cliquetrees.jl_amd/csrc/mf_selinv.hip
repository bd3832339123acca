// mf_selinv.hip — gfx950 kernel of the selected inversion of a Cholesky
// factor (selinv!, src/Multifrontal.jl/src/selinv.jl:38-424).
//
//   k_selinv   persistent: workgroups take tickets of the schedule built by
//              mf_selinv_plan.cpp (root first) and run one task each:
//              ST_XDIAG   X(K, K) = W_K, a diagonal tile of L11^-1 from the
//                         factor's inverted 64 x 64 tiles
//              ST_GEMM    one 64 x 64 tile C = sum of up to two products
//                         sgn A B' (operands plain or transposed), on MFMA:
//                         the recursion of X = L11^-1, Lt = L21 X,
//                         Sigma21 = -M Lt, Sigma11 = X' X - Lt' Sigma21
//              ST_GATHER  a column tile of M = Sigma(sep, sep) from the
//                         parent's frontal matrix (selinv_send!,
//                         copygathertri!, utils.jl:529-553), both triangles
//              ST_SMALL   a whole front with nJ <= 64 in LDS
//              (selinv_loop_snd!/nod!, selinv.jl:122-379).
// Synchronization is the tile-DAG factorization's: a task waits (bounded,
// mf_sync.h) on counters advanced only by tasks with smaller tickets; every
// value a task reads that this launch wrote is a device-coherent load and
// every output a device-coherent store; a task publishes once all its waves
// drained their stores and met at a barrier.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mf_kernels.h"
#include "mf_mfma.h"
#include "mf_sync.h"

namespace mf {

namespace {

template <typename T>
__device__ __forceinline__ T ldc(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void stc(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void drain_stores() { __builtin_amdgcn_s_waitcnt(0); }

constexpr int SKC = 32;      // k-chunk of the tile product
constexpr int SLD = 64 + 4;  // leading dimension of its LDS stages
constexpr int FLD = 64 + 1;  // leading dimension of LDS fronts and tiles

// One element of the parent's frontal matrix [Sigma11, .; Sigma21, M]
// (lower triangle, p >= q, positions in the parent's row list).
template <typename T>
__device__ __forceinline__ T parent_entry(const DevSel<T>& S, const Front& pf, int64_t pm, int p, int q) {
    const int nJ = pf.nn + pf.na;
    if (q < pf.nn) return ldc(S.base[SP_SIG] + pf.poff + p + (int64_t)q * nJ);
    return ldc(S.base[SP_M] + pm + (p - pf.nn) + (int64_t)(q - pf.nn) * pf.na);
}

// One output tile: acc = sum over the task's products of sgn A B' (wave w
// owns the 32 x 32 quarter (w & 1, w >> 1) as 2 x 2 MFMA blocks), then
// C = acc (lower triangle only for a diagonal tile).  Each k-chunk of A and
// B goes through LDS; the next chunk is fetched into registers while the
// current one is multiplied.  A plain operand is fetched down its columns
// (lane = row), a transposed one along its rows (lane = k), so both are
// coalesced.
template <typename T>
__device__ void sel_tile(const DevSel<T>& S, const SelTask& t, T* As, T* Bs, const int tid) {
    typedef typename Mfma<T>::v4 v4;
    constexpr int PER = SKC * 64 / 256;
    const int wave = tid >> 6, lane = tid & 63;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    v4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4{0, 0, 0, 0};
    for (int q = 0; q < t.np; ++q) {
        const SelPart p = t.p[q];
        const T* A = S.base[p.arrA] + p.a_off;
        const T* B = S.base[p.arrB] + p.b_off;
        const T sg = (T)p.sgn;
        T ra[PER], rb[PER];
        // element i of this thread in a chunk: plain (row tid & 63, k (tid >> 6) + 4 i),
        // transposed (row (tid >> 5) + 8 i, k tid & 31)
        auto fetch = [&](int k0) {
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                int r, gk;
                if (p.ta) { r = (tid >> 5) + 8 * i; gk = k0 + (tid & 31); }
                else { r = tid & 63; gk = k0 + (tid >> 6) + 4 * i; }
                T v = (T)0;
                if (r < t.m && gk < p.k) v = ldc(p.ta ? A + gk + (int64_t)r * p.lda : A + r + (int64_t)gk * p.lda);
                ra[i] = v;
                if (p.tb) { r = (tid >> 5) + 8 * i; gk = k0 + (tid & 31); }
                else { r = tid & 63; gk = k0 + (tid >> 6) + 4 * i; }
                v = (T)0;
                if (r < t.n && gk < p.k) v = ldc(p.tb ? B + gk + (int64_t)r * p.ldb : B + r + (int64_t)gk * p.ldb);
                rb[i] = v;
            }
        };
        fetch(0);
        for (int k0 = 0; k0 < p.k; k0 += SKC) {
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int ai = p.ta ? (tid & 31) * SLD + (tid >> 5) + 8 * i : ((tid >> 6) + 4 * i) * SLD + (tid & 63);
                const int bi = p.tb ? (tid & 31) * SLD + (tid >> 5) + 8 * i : ((tid >> 6) + 4 * i) * SLD + (tid & 63);
                As[ai] = sg * ra[i];
                Bs[bi] = rb[i];
            }
            __syncthreads();
            if (k0 + SKC < p.k) fetch(k0 + SKC);
#pragma unroll
            for (int kk = 0; kk < SKC; kk += 4) {
                const T a0 = As[(kk + lk) * SLD + wr + li];
                const T a1 = As[(kk + lk) * SLD + wr + 16 + li];
                const T b0 = Bs[(kk + lk) * SLD + wc + li];
                const T b1 = Bs[(kk + lk) * SLD + wc + 16 + li];
                acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
                acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
                acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
                acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
            }
            __syncthreads();
        }
    }
    T* C = S.base[t.arrC] + t.c_off;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wr + i * 16 + Mfma<T>::row(lk, v), col = wc + j * 16 + li;
                if (row < t.m && col < t.n && (!t.lower || row >= col))
                    stc(C + row + (int64_t)col * t.ldc, acc[i][j][v]);
            }
}

// X(K, K) = W_K (kb x kb, zeros above the diagonal included).
template <typename T>
__device__ void sel_xdiag(const DevPlan& P, const DevSel<T>& S, int J, int K, const int tid) {
    const Front f = P.fronts[J];
    const int kb = min(NB, f.nn - NB * K);
    const T* W = S.base[SP_W] + f.woff + (int64_t)NB * K;
    T* X = S.base[SP_X] + S.xoff[J] + (int64_t)NB * K + (int64_t)NB * K * f.nn;
    const int r = tid & 63;
    for (int c = tid >> 6; c < kb; c += 4)
        if (r < kb) stc(X + r + (int64_t)c * f.nn, ldc(W + r + (int64_t)c * f.nn));
}

// Column tile jb of M (rows jb*64 .. na-1 and, by symmetry, the same tiles
// above the diagonal): M(a, b) = F_parent(rel[max(a, b)], rel[min(a, b)]).
template <typename T>
__device__ void sel_gather(const DevPlan& P, const DevSel<T>& S, int J, int jb, T* Ts, const int tid) {
    const Front f = P.fronts[J];
    const Front pf = P.fronts[f.parent];
    const int64_t pm = S.moff[f.parent];
    const int32_t* rel = P.rel + f.reloff;
    T* M = S.base[SP_M] + S.moff[J];
    const int na = f.na, b0 = NB * jb, cb = min(NB, na - b0);
    for (int a0 = b0; a0 < na; a0 += NB) {
        const int rb = min(NB, na - a0);
        for (int e = tid; e < NB * NB; e += 256) {
            const int r = e & 63, c = e >> 6;
            if (r < rb && c < cb) {
                const int a = a0 + r, b = b0 + c;
                const int pa = rel[max(a, b)], pb = rel[min(a, b)];
                const T v = parent_entry(S, pf, pm, pa, pb);
                stc(M + a + (int64_t)b * na, v);
                Ts[r + c * FLD] = v;
            }
        }
        __syncthreads();
        if (a0 != b0) {  // the mirrored tile, written down its columns
            for (int e = tid; e < NB * NB; e += 256) {
                const int x = e & 63, y = e >> 6;
                if (x < cb && y < rb) stc(M + (b0 + x) + (int64_t)(a0 + y) * na, Ts[y + x * FLD]);
            }
        }
        __syncthreads();
    }
}

// A whole front with nJ <= 64 in LDS: L = [L11; L21] in Lf (ld FLD), M in
// Mf; Lt = L21 L11^-1 and X = L11^-1 in place, Sigma21 = -M Lt into Mf,
// Sigma11 = X' X - Lt' Sigma21 straight to the output panel.
template <typename T>
__device__ void sel_small(const DevPlan& P, const DevSel<T>& S, int J, T* Lf, T* Mf, const int tid) {
    const Front f = P.fronts[J];
    const int nn = f.nn, na = f.na, nJ = nn + na;
    const T* Lp = S.base[SP_PANEL] + f.poff;
    for (int e = tid; e < nJ * nn; e += 256) {
        const int r = e % nJ, c = e / nJ;
        Lf[r + c * FLD] = Lp[r + (int64_t)c * nJ];
    }
    if (na > 0) {
        const Front pf = P.fronts[f.parent];
        const int64_t pm = S.moff[f.parent];
        const int32_t* rel = P.rel + f.reloff;
        T* Mo = S.base[SP_M] + S.moff[J];
        const bool keep = f.chd1 > f.chd0;  // the children gather from it
        for (int e = tid; e < na * na; e += 256) {
            const int a = e % na, b = e / na;
            const T v = parent_entry(S, pf, pm, rel[max(a, b)], rel[min(a, b)]);
            Mf[a + b * FLD] = v;
            if (keep) stc(Mo + a + (int64_t)b * na, v);
        }
    }
    __syncthreads();
    // Lt = L21 L11^-1, last column first: Lt(:, c) = L21'(:, c) / L(c, c), then
    // L21'(:, j) -= Lt(:, c) L(c, j) for j < c
    for (int c = nn - 1; c >= 0 && na > 0; --c) {
        const T inv = (T)1 / Lf[c + c * FLD];
        for (int r = tid; r < na; r += 256) Lf[nn + r + c * FLD] *= inv;
        __syncthreads();
        for (int e = tid; e < na * c; e += 256) {
            const int r = e % na, j = e / na;
            Lf[nn + r + j * FLD] -= Lf[nn + r + c * FLD] * Lf[c + j * FLD];
        }
        __syncthreads();
    }
    // X = L11^-1 in place, last column first: X(i, j) = -X(j, j) sum_{k=j+1}^{i} X(i, k) L(k, j)
    for (int j = nn - 1; j >= 0; --j) {
        const T xjj = (T)1 / Lf[j + j * FLD];
        T v = (T)0;
        const int i = j + 1 + tid;
        if (i < nn) {
            T s = (T)0;
            for (int k = j + 1; k <= i; ++k) s += Lf[i + k * FLD] * Lf[k + j * FLD];
            v = -xjj * s;
        }
        __syncthreads();
        if (i < nn) Lf[i + j * FLD] = v;
        if (tid == 0) Lf[j + j * FLD] = xjj;
        __syncthreads();
    }
    // Sigma21 = -M Lt (na x nn, at most 4 entries per thread), then into Mf
    T s21[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = tid + 256 * q;
        T s = (T)0;
        if (e < na * nn) {
            const int r = e % na, c = e / na;
            for (int k = 0; k < na; ++k) s += Mf[r + k * FLD] * Lf[nn + k + c * FLD];
        }
        s21[q] = -s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = tid + 256 * q;
        if (e < na * nn) Mf[e % na + (e / na) * FLD] = s21[q];
    }
    __syncthreads();
    T* Sg = S.base[SP_SIG] + f.poff;
    for (int e = tid; e < nJ * nn; e += 256) {
        const int r = e % nJ, c = e / nJ;
        T v = (T)0;
        if (r >= nn) {
            v = Mf[(r - nn) + c * FLD];
        } else if (r >= c) {
            for (int k = r; k < nn; ++k) v += Lf[k + r * FLD] * Lf[k + c * FLD];
            T s = (T)0;
            for (int k = 0; k < na; ++k) s += Lf[nn + k + r * FLD] * Mf[k + c * FLD];
            v -= s;
        }
        stc(Sg + r + (int64_t)c * nJ, v);
    }
}

template <typename T>
__global__ __launch_bounds__(256, 2) void k_selinv(DevPlan P, DevSel<T> S) {
    constexpr int SM1 = 2 * 64 * FLD, SM2 = 2 * SKC * SLD;
    __shared__ T smem[SM1 > SM2 ? SM1 : SM2];
    __shared__ int s_t;
    for (;;) {
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        if (tid == 0) s_t = (int)atomicAdd(S.ticket, 1u);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_t);
        __syncthreads();
        if (t >= S.ntask) break;
        const SelTask tk = S.task[t];
        if (tid == 0)
            for (int w = 0; w < 2; ++w)
                if (tk.wait[w] >= 0) (void)bounded_wait(S.cnt + tk.wait[w], tk.target[w], S.hang, S.ticket, S.stall, 1);
        __syncthreads();
        switch (tk.kind) {
            case ST_XDIAG: sel_xdiag<T>(P, S, tk.front, tk.a, tid); break;
            case ST_GEMM: sel_tile<T>(S, tk, smem, smem + SKC * SLD, tid); break;
            case ST_GATHER: sel_gather<T>(P, S, tk.front, tk.a, smem, tid); break;
            default: sel_small<T>(P, S, tk.front, smem, smem + 64 * FLD, tid); break;
        }
        drain_stores();
        __syncthreads();
        if (tid == 0) {
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_fetch_add(S.cnt + tk.front, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // closes the thread-0 block before the back-edge
        __syncthreads();
    }
}

}  // namespace

template <typename T>
void launch_selinv(const DevPlan& P, const DevSel<T>& S, hipStream_t s) {
    if (S.ntask <= 0) return;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_selinv<T>, 256, 0);
    int grid = std::max(1, cus) * std::max(1, per);
    if (grid > S.ntask) grid = S.ntask;
    hipLaunchKernelGGL(k_selinv<T>, dim3(grid), dim3(256), 0, s, P, S);
}

template void launch_selinv<double>(const DevPlan&, const DevSel<double>&, hipStream_t);
template void launch_selinv<float>(const DevPlan&, const DevSel<float>&, hipStream_t);

}  // namespace mf
