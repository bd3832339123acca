// mf_factor.hip — gfx950 kernels of the multifrontal numeric factorization.
//
//   k_small     fused front kernel for internal fronts with nJ <= 64: the
//               front lives in LDS; extend-add of the children's update
//               matrices, assembly, partial factorization and the write-out
//               of the panel and of U in one workgroup.  Persistent mode
//               hands out fronts by ticket in postorder levels and waits on
//               the children's completion flags (bounded wait).
//               Reference: chol_loop_snd!/nod! cholesky.jl:434-692,
//               chol_send! cholesky.jl:950-983, potrf2! blas/potrf.jl:3-41.
//   k_dag       persistent tile-DAG over all big fronts: extend-add column
//               groups (F11/F21 += F, cholesky.jl:520-521; M22 <- F22,
//               cholesky.jl:542), small fronts above the small region,
//               diagonal-tile factor + inverse (potrf2!), panel solve tiles
//               (trsm!, blas/trsx.jl:156-159, as a product with the inverted
//               diagonal tile), trailing-update tiles and update-matrix tiles
//               M22 -= L21 [D] L21' (syrk!, blas/syrk.jl:137-159).
//   k_diag      diagonal-tile inverse alone (solves of a factor given in the
//               reference layout).
//   k_copyout   internal panels -> the reference's Dval/Lval/d layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mf_kernels.h"
#include "mf_mfma.h"
#include "mf_sync.h"

namespace mf {

namespace {

template <typename T>
__device__ __forceinline__ double sign_of(const DevFactor<T>& F, int gcol) {
    if (F.signs) return F.signs[gcol];
    return F.diag == 0 ? 1.0 : 0.0;
}

// regularize(R, S, Djj, j): no_regularization.jl:27-33 and
// dynamic_regularization.jl:75-90.  0 signals a failed pivot.
template <typename T>
__device__ __forceinline__ double regularize(const DevFactor<T>& F, double s, double Djj) {
    if (F.reg_kind == 1) return (s * Djj < F.eps) ? F.delta * s : Djj;
    if (s != 0.0 && !(s * Djj > 0.0)) return 0.0;
    return Djj;
}

// Julia's max of floats: NaN when either argument is NaN.
__device__ __forceinline__ double jmax(double a, double b) {
    if (a != a || b != b) return a + b;
    return a > b ? a : b;
}

// GMW81 regularize (gmw81.jl:109-155): s * max(|Djj|, amax^2 / beta, delta),
// amax the largest magnitude below the pivot in its front column as the
// column stands after the previous pivots (GMW81 never fails a pivot).
template <typename T>
__device__ __forceinline__ double regularize_gmw(const DevFactor<T>& F, double s, double Djj, double amax) {
    return s * jmax(jmax(fabs(Djj), amax * amax / F.gmw[2]), F.delta);
}

// Largest value over the wavefront (NaN-propagating), in every lane.
__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = jmax(v, __shfl_xor(v, o));
    return v;
}

// Waits until this wave's global stores (and loads) have completed.  Task
// outputs are device-coherent stores, so once every wave has drained before
// the closing barrier, the counter update that thread 0 makes afterwards
// follows all of the task's data.
__device__ __forceinline__ void drain_stores() { __builtin_amdgcn_s_waitcnt(0); }

// Bounded wait until *flag >= target (the watchdog of mf_sync.h).
template <typename T>
__device__ __forceinline__ void wait_ge(int32_t* flag, int32_t target, const DevFactor<T>& F) {
    (void)bounded_wait(flag, target, F.hang, F.ticket, F.stall, 2);
}

__device__ __forceinline__ int ld_rel(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Update matrices of small fronts are written and read with device-coherent
// (sc1) accesses: a small front then publishes with its stores drained and a
// flag, and its parent reads U without a cache-wide invalidate.
template <typename T>
__device__ __forceinline__ T ldc(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void stc(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int LDS = SMALL_NJ + 1;  // leading dimension of LDS fronts and tiles
constexpr int KC = 32;             // k-chunk of the GEMM tile
constexpr int LDT = 64 + 4;        // leading dimension of the GEMM LDS stages

// Right-looking partial factorization of the first nn columns of the
// nj x nj front M (lower triangle, leading dimension LDS) in LDS.
// Returns after a failed pivot (recorded in F.info).
template <typename T>
__device__ __forceinline__ void factor_lds(T* M, int nn, int nj, int gc0, const DevFactor<T>& F,
                           T* s_piv, T* s_inv, int* s_fail, const int tid) {
    for (int k = 0; k < nn; ++k) {
        double amax = 0.0;
        if (F.reg_kind == 2 && tid < 64)  // GMW81: the column below the pivot (nj <= 64 rows)
            amax = wave_max((tid > k && tid < nj) ? fabs((double)M[tid + k * LDS]) : 0.0);
        if (tid == 0) {
            double akk = (double)M[k + k * LDS];
            double piv = F.reg_kind == 2 ? regularize_gmw(F, sign_of(F, gc0 + k), akk, amax)
                                         : regularize(F, sign_of(F, gc0 + k), akk);
            if (piv == 0.0) {
                *s_fail = 1;
                atomicMin(F.info, (unsigned long long)(gc0 + k + 1));
            } else if (F.diag == 0) {
                T r = (T)sqrt(piv);
                M[k + k * LDS] = r;
                *s_piv = (T)1;
                *s_inv = (T)1 / r;
            } else {
                stc(F.d + gc0 + k, (T)piv);
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

// One small front (nJ <= 64) in LDS: extend-add of the children's update
// matrices, assembly with the panel's A entries, partial factorization,
// write-out of the panel and of U.  The caller has made the children final.
template <typename T>
__device__ __forceinline__ void small_front(const DevPlan& P, const DevFactor<T>& F, int J, T* M, T* s_piv,
                            T* s_inv, int* s_fail, const int tid) {
    const Front f = P.fronts[J];
    const int nn = f.nn, na = f.na, nj = nn + na;
    const int r = tid & 63, cg = tid >> 6;
    for (int c = cg; c < nj; c += 4) M[r + c * LDS] = (T)0;
    __syncthreads();
    for (int q = f.chd0; q < f.chd1; ++q) {
        const Front kf = P.fronts[P.chd[q]];
        const int ma = kf.na;
        const int32_t* relK = P.rel + kf.reloff;
        const T* UK = F.U + kf.uoff;
        if (r < ma) {
            // all loads first (fixed count, clamped in-bounds addresses), then
            // the adds: loads under a per-lane loop bound would each be waited
            // for inside the loop
            const int rr = relK[r];
            for (int i0 = 0; i0 < SMALL_NJ / 4; i0 += 8) {
                if (cg + 4 * i0 > r) break;
                int cc[8];
                T v[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int c = cg + 4 * (i0 + i), cs = c <= r ? c : 0;
                    cc[i] = relK[cs];
                    v[i] = ldc(UK + r + cs * ma);
                }
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (cg + 4 * (i0 + i) <= r) M[rr + cc[i] * LDS] += v[i];
            }
        }
        __syncthreads();
    }
    T* Pj = F.panel + f.poff;
    if (r < nj) {
        for (int i0 = 0; i0 < SMALL_NJ / 4; i0 += 8) {
            if (cg + 4 * i0 > r || cg + 4 * i0 >= nn) break;
            T v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int c = cg + 4 * (i0 + i);
                v[i] = (c <= r && c < nn) ? Pj[r + c * nj] : (T)0;
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int c = cg + 4 * (i0 + i);
                if (c <= r && c < nn) M[r + c * LDS] += v[i];
            }
        }
    }
    __syncthreads();
    factor_lds<T>(M, nn, nj, f.c0, F, s_piv, s_inv, s_fail, tid);
    if (r < nj)
        for (int c = cg; c <= r && c < nn; c += 4) stc(Pj + r + c * nj, M[r + c * LDS]);
    if (na > 0 && r < na) {
        T* Uj = F.U + f.uoff;
        for (int c = cg; c <= r; c += 4) stc(Uj + r + c * na, M[(nn + r) + (nn + c) * LDS]);
    }
}

constexpr int LDT1 = TINY_NJ + 1;     // leading dimension of a wavefront's LDS front

// One tiny front (nJ <= TINY_NJ) by one wavefront in its own LDS slab Mw:
// the same extend-add, assembly, right-looking partial factorization and
// write-out as small_front, lane r keeping row r.  A wavefront's LDS
// operations complete in order, so no barrier is needed.
template <typename T>
__device__ __forceinline__ void wave_front(const DevPlan& P, const DevFactor<T>& F, int J, T* Mw,
                                           const int lane) {
    const Front f = P.fronts[J];
    const int nn = f.nn, na = f.na, nj = nn + na, gc0 = f.c0;
    const int r = lane;
    if (r < nj)
        for (int c = 0; c <= r; ++c) Mw[r + c * LDT1] = (T)0;
    for (int q = f.chd0; q < f.chd1; ++q) {
        const Front kf = P.fronts[P.chd[q]];
        const int ma = kf.na;
        const int32_t* relK = P.rel + kf.reloff;
        const T* UK = F.U + kf.uoff;
        if (r < ma) {
            // loads first, eight at a time (fixed count, clamped in-bounds
            // addresses), then the adds
            const int rr = relK[r];
#pragma unroll
            for (int c0 = 0; c0 < TINY_NJ; c0 += 8) {
                if (c0 > r) break;
                int cc[8];
                T v[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int c = c0 + i, cs = c <= r ? c : 0;
                    cc[i] = relK[cs];
                    v[i] = ldc(UK + r + cs * ma);
                }
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (c0 + i <= r) Mw[rr + cc[i] * LDT1] += v[i];
            }
        }
    }
    T* Pj = F.panel + f.poff;
    if (r < nj) {
#pragma unroll
        for (int c0 = 0; c0 < TINY_NJ; c0 += 8) {
            if (c0 > r || c0 >= nn) break;
            T v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int c = c0 + i;
                v[i] = (c <= r && c < nn) ? Pj[r + c * nj] : (T)0;
            }
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (c0 + i <= r && c0 + i < nn) Mw[r + (c0 + i) * LDT1] += v[i];
        }
    }
    for (int k = 0; k < nn; ++k) {
        const T akk = Mw[k + k * LDT1];
        double piv;
        if (F.reg_kind == 2) {  // GMW81: the column below the pivot, one row per lane
            const double amax = wave_max((r > k && r < nj) ? fabs((double)Mw[r + k * LDT1]) : 0.0);
            piv = regularize_gmw(F, sign_of(F, gc0 + k), (double)akk, amax);
        } else {
            piv = regularize(F, sign_of(F, gc0 + k), (double)akk);
        }
        if (piv == 0.0) {
            if (lane == 0) atomicMin(F.info, (unsigned long long)(gc0 + k + 1));
            return;  // the front stays unfinished; info reports the column
        }
        T inv, spv;
        if (F.diag == 0) {
            const T rt = (T)sqrt(piv);
            if (lane == 0) Mw[k + k * LDT1] = rt;
            inv = (T)1 / rt;
            spv = (T)1;
        } else {
            if (lane == 0) stc(F.d + gc0 + k, (T)piv);
            inv = (T)1 / (T)piv;
            spv = (T)piv;
        }
        if (r > k && r < nj) Mw[r + k * LDT1] *= inv;
        if (r > k && r < nj) {
            const T lr = Mw[r + k * LDT1] * spv;
            for (int c = k + 1; c <= r; ++c) Mw[r + c * LDT1] -= lr * Mw[c + k * LDT1];
        }
    }
    if (r < nj)
        for (int c = 0; c <= r && c < nn; ++c) stc(Pj + r + c * nj, Mw[r + c * LDT1]);
    if (na > 0 && r < na) {
        T* Uj = F.U + f.uoff;
        for (int c = 0; c <= r; ++c) stc(Uj + r + c * na, Mw[(nn + r) + (nn + c) * LDT1]);
    }
}

// Persistent walk of the small region.  Tickets are units of the planner's
// children-first order: one front of nJ > TINY_NJ (the whole workgroup), or
// up to four consecutive tiny fronts (one per wavefront, each waiting for
// its own children; a child in the same unit belongs to another wavefront,
// which progresses on its own).  Update matrices go through device-coherent
// accesses, so a front is published by draining its stores and setting its
// flag; no cache-wide write-back or invalidate.
template <typename T>
__global__ __launch_bounds__(256, 4) void k_small(DevPlan P, DevFactor<T> F, const int32_t* list,
                                               const int32_t* units, int count) {
    constexpr int MS = SMALL_NJ * LDS > 4 * TINY_NJ * LDT1 ? SMALL_NJ * LDS : 4 * TINY_NJ * LDT1;
    __shared__ T M[MS];
    __shared__ T s_piv, s_inv;
    __shared__ int s_fail, s_t;
    for (;;) {
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        if (tid == 0) s_t = (int)atomicAdd(F.ticket, 1u);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= count) break;
        const int u = units[t], start = u >> 3, cnt = u & 7;
        if (cnt == 0) {  // one front, the whole workgroup
            const int J = list[start];
            if (tid == 0) {
                s_fail = 0;
                const Front f = P.fronts[J];
                for (int q = f.chd0; q < f.chd1; ++q) wait_ge(F.flags + P.chd[q], 1, F);
            }
            __syncthreads();
            small_front<T>(P, F, J, M, &s_piv, &s_inv, &s_fail, tid);
            drain_stores();
            __syncthreads();
            if (tid == 0) __hip_atomic_store(F.flags + J, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const int wave = tid >> 6, lane = tid & 63;
            if (wave < cnt) {
                const int J = list[start + wave];
                const Front f = P.fronts[J];
                if (lane == 0)
                    for (int q = f.chd0; q < f.chd1; ++q) wait_ge(F.flags + P.chd[q], 1, F);
                // the children's data is read after the wait (device-coherent loads)
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                wave_front<T>(P, F, J, M + wave * TINY_NJ * LDT1, lane);
                drain_stores();
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) __hip_atomic_store(F.flags + J, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        // closes the ticket before the back-edge: every thread reaches the
        // next ticket together
        __syncthreads();
    }
}

// Extend-add of one parent column by one wavefront: for each window of
// EA_WIN parent rows the children's contributions (in the reference's
// reverse child order) are accumulated in a wave-private LDS window, then
// added to the panel (F11/F21 += F, cholesky.jl:520-521) or written to U
// (M22 <- F22, cholesky.jl:542).  The child rows that land in window w are
// [wstart[w], wstart[w+1]) (rel is increasing), so no search is needed; a
// wavefront's LDS operations complete in order, so no barrier is needed.
// The column's destination and its sources' addresses come precomputed
// (EaDst, EaSrc; source s in lane s % 64), the panel values of the window
// are fetched together with the window bounds, and the next source's rows
// are fetched while the current source's are added.  A wavefront takes the
// columns e_first, e_first + 4, ... < e_end; the records of the next column
// are fetched while the current one is processed.
// acc[i] += v in LDS without returning the old value: no wait for a read
// before the add (a read-add-write per element is one LDS round trip each,
// since the compiler cannot tell the indices apart).  A wavefront's LDS
// operations complete in order, so the additions happen in program order.
template <typename T>
__device__ __forceinline__ void lds_add(T* p, T v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

template <typename T>
__device__ __forceinline__ void ea_columns(const DevPlan& P, const DevFactor<T>& F, int e_first, int e_end,
                                           T* acc, int lane) {
    EaTask e_nx = P.ea[e_first];
    EaDst d_nx = P.ea_dst[e_first];
    EaSrc s_nx{};
    if (e_nx.src1 - e_nx.src0 <= 64 && lane < e_nx.src1 - e_nx.src0) s_nx = P.ea_srcx[e_nx.src0 + lane];
    for (int ei = e_first; ei < e_end; ei += 4) {
        const EaTask e = e_nx;
        const EaDst ed = d_nx;
        int64_t my_u = s_nx.u_off;
        int my_rel = s_nx.rel_off, my_ws = s_nx.ws_off, my_tcol = s_nx.tcol, my_f0 = s_nx.i0, my_f1 = s_nx.i1;
        if (ei + 4 < e_end) {
            e_nx = P.ea[ei + 4];
            d_nx = P.ea_dst[ei + 4];
            const int nsn = e_nx.src1 - e_nx.src0;
            if (nsn <= 64 && lane < nsn) s_nx = P.ea_srcx[e_nx.src0 + lane];
        }
        T* dst = (ed.to_u ? F.U : F.panel) + ed.off;  // dst[p]: parent row p
        const bool to_u = ed.to_u != 0;
        const int ns = e.src1 - e.src0;
        auto load_meta = [&](int sb) {
            if (sb + lane < ns) {
                const EaSrc x = P.ea_srcx[e.src0 + sb + lane];
                my_u = x.u_off;
                my_rel = x.rel_off;
                my_ws = x.ws_off;
                my_tcol = x.tcol;
                my_f0 = x.i0;
                my_f1 = x.i1;
            }
        };
        const int wfirst = e.row0 / EA_WIN;
        for (int w0 = (e.row0 / EA_WIN) * EA_WIN; w0 < e.row1; w0 += EA_WIN) {
            const int w = w0 / EA_WIN;
            T pre[EA_WIN / 64];
#pragma unroll
            for (int j = 0; j < EA_WIN / 64; ++j) {
                const int p = w0 + lane + 64 * j;
                pre[j] = (!to_u && p >= e.row0 && p < e.row1) ? dst[p] : (T)0;
                acc[lane + 64 * j] = (T)0;
            }
            for (int sb = 0; sb < ns; sb += 64) {
                if (ns > 64) load_meta(sb);
                const int nsb = min(64, ns - sb);
                int my_i0 = my_f0, my_i1 = my_f1;  // the first window's bounds come with the record
                if (lane < nsb && w != wfirst) {
                    const int32_t* ws = P.wstart + my_ws;
                    my_i0 = max(ws[w], my_tcol);
                    my_i1 = ws[w + 1];
                }
                // rows i0 + lane + 64 j (j < 4) of source q.  The loads are
                // unconditional (rows past the source's end read its row 0,
                // always in bounds, and are flagged in ok), so the eight loads
                // of a fetch are in flight together; a load under a per-lane
                // branch would be waited for inside that branch.
                int rr[4];
                T vv[4];
                unsigned ok = 0;
                auto fetch = [&](int q, int i) {
                    const int i1 = __builtin_amdgcn_readlane(my_i1, q);
                    const int relo = __builtin_amdgcn_readlane(my_rel, q);
                    const unsigned ulo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(uint64_t)my_u, q);
                    const unsigned uhi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((uint64_t)my_u >> 32), q);
                    const int64_t uo = (int64_t)(((uint64_t)uhi << 32) | ulo);
                    ok = 0;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int ii = i + lane + 64 * j;
                        const int ic = ii < i1 ? ii : 0;
                        ok |= (ii < i1 ? 1u : 0u) << j;
                        rr[j] = P.rel[relo + ic];
                        vv[j] = ldc(F.U + uo + ic);
                    }
                };
                if (nsb > 0) fetch(0, __builtin_amdgcn_readlane(my_i0, 0));
                for (int q = 0; q < nsb; ++q) {
                    const int i0 = __builtin_amdgcn_readlane(my_i0, q), i1 = __builtin_amdgcn_readlane(my_i1, q);
                    int cr[4];
                    T cv[4];
                    const unsigned cok = ok;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        cr[j] = rr[j];
                        cv[j] = vv[j];
                    }
                    if (q + 1 < nsb) fetch(q + 1, __builtin_amdgcn_readlane(my_i0, q + 1));
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if ((cok >> j) & 1u) lds_add(acc + (cr[j] - w0), cv[j]);
                    // rows beyond the first 256 of this source in the window
                    for (int i = i0 + 256; i < i1; i += 256) {
                        const int relo = __builtin_amdgcn_readlane(my_rel, q);
                        const unsigned ulo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(uint64_t)my_u, q);
                        const unsigned uhi =
                            (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((uint64_t)my_u >> 32), q);
                        const int64_t uo = (int64_t)(((uint64_t)uhi << 32) | ulo);
                        int r2[4];
                        T v2[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int ii = i + lane + 64 * j;
                            const int ic = ii < i1 ? ii : 0;
                            r2[j] = P.rel[relo + ic];
                            v2[j] = ldc(F.U + uo + ic);
                        }
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (i + lane + 64 * j < i1) lds_add(acc + (r2[j] - w0), v2[j]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < EA_WIN / 64; ++j) {
                const int p = w0 + lane + 64 * j;
                if (p >= e.row0 && p < e.row1) stc(dst + p, to_u ? acc[lane + 64 * j] : pre[j] + acc[lane + 64 * j]);
            }
        }
    }
}

constexpr int PB = 16;                          // block size inside the diagonal tile
constexpr int PTILE = NB * LDS + 3 * PB * PB + 2 * NB + PB;  // LDS elements of potrf_blocked

// Value of v held by lane l of the wavefront (l wave-uniform).
__device__ __forceinline__ double bcast(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float bcast(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// 16 x 16 block product on the matrix cores: acc += sum over the K = 16 * kt
// columns of  opA(i, k) * opB(k, j), with opA(i, k) = X[(ra + i) + (ca + k) * LDS]
// (times sc[ca + k] when sc is given) and opB(k, j) = Y[(rb + k) + (cb + j) * ldy].
template <typename T>
__device__ __forceinline__ void blk_mma(const T* X, int ra, int ca, const T* Y, int rb, int cb, int ldy,
                                        int kt, const T* sc, typename Mfma<T>::v4& acc) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    for (int s = 0; s < 4 * kt; ++s) {
        const int k = 4 * s + lk;
        T av = X[(ra + li) + (ca + k) * LDS];
        if (sc) av *= sc[ca + k];
        acc = Mfma<T>::op(av, Y[(rb + k) + (cb + li) * ldy], acc);
    }
}

// Factors the 16 x 16 diagonal block held by lanes 0..15 of one wavefront
// (lane i: row i in x[0..15]; the other lanes compute alongside and are
// ignored), nk <= 16 columns.  Per pivot the pivot itself comes by readlane
// (a short chain to the next pivot), and column k reaches every lane through
// LDS (cb) for the rank-1 update.  pv/rd receive the pivots and the
// reciprocals of L's diagonal (Cholesky) or of the pivots (LDLt).  RULE 0: the general pivot rule (signs,
// regularization: regularize per pivot); RULE 1: Cholesky with the default
// signs (fails unless the pivot is positive); RULE 2: LDLt with the default
// signs (fails on a zero pivot only).
template <typename T, int RULE>
__device__ __forceinline__ void pivot_block(T (&x)[PB], const DevFactor<T>& F, const int gcol0, const int nk,
                                            const bool chol, T* cb, T* pv, T* rd, const int lane) {
    // the pivot of the next step, as it will stand after this step's update:
    // lane k+1 forms it from its own entries (x[k+1] - ar * ark is the same
    // fused multiply-add as the rank-1 update below applies to its x[k+1],
    // since the broadcast column entry of row k+1 is its own x[k]), so the
    // chain from pivot to pivot skips the LDS round trip of the column
    T dnext = bcast(x[0], 0);
#pragma unroll
    for (int k = 0; k < PB; ++k) {
        if (k >= nk) continue;  // uniform
        // the lane id is re-read per step so that the compiler does not keep
        // sixteen steps' worth of lane masks live in scalar registers
        int ln = lane;
        asm volatile("" : "+v"(ln));
        // column k as it stands to every lane through LDS: one write, then
        // uniform reads (a wavefront's LDS operations complete in order)
        if (ln < PB) cb[ln] = x[k];
        __builtin_amdgcn_wave_barrier();
        const T akk = dnext;
        double piv;
        bool bad;
        if (RULE == 0) {
            piv = regularize(F, sign_of(F, gcol0 + k), (double)akk);
            bad = (piv == 0.0);
        } else {
            piv = (double)akk;
            bad = RULE == 1 ? !(piv > 0.0) : (piv == 0.0);
        }
        if (bad) piv = 1.0;  // keeps the rest finite after a failure
        // 1/sqrt(piv) and 1/piv from the hardware estimates plus two Newton
        // steps each (full double precision, a short dependency chain)
        T ip, sq = (T)1, isq = (T)1;
        if (chol) {
            double y = __builtin_amdgcn_rsq(piv);
            const double h = 0.5 * piv;
            y = y * __builtin_fma(-h * y, y, 1.5);
            y = y * __builtin_fma(-h * y, y, 1.5);
            isq = (T)y;
            sq = (T)(piv * y);
            ip = (T)(y * y);
        } else {
            double y = __builtin_amdgcn_rcp(piv);
            y = __builtin_fma(__builtin_fma(-piv, y, 1.0), y, y);
            y = __builtin_fma(__builtin_fma(-piv, y, 1.0), y, y);
            ip = (T)y;
        }
        T colv[PB];
#pragma unroll
        for (int j = k + 1; j < PB; ++j) colv[j] = cb[j];
        const T ark = x[k], ar = ark * ip;
        if (k + 1 < PB) dnext = bcast(x[k + 1] - ar * ark, k + 1);
        if (ln == 0) {
            if (bad) atomicMin(F.info, (unsigned long long)(gcol0 + k + 1));
            else if (!chol) stc(F.d + gcol0 + k, (T)piv);
            pv[k] = (T)piv;
            rd[k] = chol ? isq : ip;
        }
#pragma unroll
        for (int j = k + 1; j < PB; ++j) x[j] -= ar * colv[j];
        if (ln >= k) {
            if (chol) {
                x[k] = (ln == k) ? sq : ark * isq;
            } else if (ln > k) {
                x[k] = ar;  // LDLt keeps the unregularized pivot on the diagonal
            }
        }
    }
}

// Diagonal tile, blocked: the 64 x 64 tile lives in LDS and is factored in
// four 16-column block steps (potrf2!'s right-looking recursion cut at 16):
// (1) one wavefront factors the 16 x 16 diagonal block in registers with the
// pivot rule (columns broadcast with readlane, no LDS round trip) and keeps
// the reciprocals of the diagonal, (2) one lane per row solves the rows below
// against it, (3) the rank-16 update of the trailing lower triangle runs on
// the matrix cores, one 16 x 16 block per wavefront.  Then W = L^-1 in place:
// the four diagonal blocks are inverted one per wavefront, and the block rows
// below follow in three stages of W_ij = -W_ii (sum_k L_ik W_kj), again on the
// matrix cores.  L goes to the panel and W (ld = nn, zeros above the
// diagonal) to the W pool.  LDS: PTILE elements.
template <typename T>
__device__ __forceinline__ void potrf_blocked(const DevFactor<T>& F, const Front& f, int k0, bool invert_only,
                              T* lds, const int tid) {
    typedef typename Mfma<T>::v4 v4;
    T* A = lds;                       // 64 x LDS, column-major
    T* Tm = lds + NB * LDS;           // 3 blocks of 16 x 16 (stage products, ld 16)
    T* pv = Tm + 3 * PB * PB;         // pivots of the tile (LDLt scaling)
    T* rd = pv + NB;                  // reciprocals of L's diagonal (Cholesky)
    T* cb = rd + NB;                  // the pivot column of the 16 x 16 block (broadcast)
    const int nj = f.nn + f.na, kb = min(NB, f.nn - k0), gc0 = f.c0 + k0;
    T* blk = F.panel + f.poff + k0 + (int64_t)k0 * nj;
    const int r = tid & 63, cg = tid >> 6, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const bool chol = F.diag == 0;
    {
        // all sixteen loads in flight (outside the tile: the tile's first
        // element stands in), then the LDS writes
        T tv[NB / 4];
#pragma unroll
        for (int i = 0; i < NB / 4; ++i) {
            const int c = cg + 4 * i;
            tv[i] = ldc(blk + ((r < kb && c < kb && c <= r) ? r + (int64_t)c * nj : 0));
        }
#pragma unroll
        for (int i = 0; i < NB / 4; ++i) {
            const int c = cg + 4 * i;
            A[r + c * LDS] = (r < kb && c < kb && c <= r) ? tv[i] : (r == c ? (T)1 : (T)0);
        }
    }
    if (tid < NB) {
        pv[tid] = (T)1;
        rd[tid] = (T)1;
    }
    __syncthreads();
    if (!invert_only) {
        for (int b = 0; b < NB / PB; ++b) {
            const int c0 = b * PB;
            if (c0 >= kb) break;  // uniform
            // (1) diagonal block by wavefront 0: lane i < 16 holds row c0 + i
            if (tid < 64) {
                T x[PB];
#pragma unroll
                for (int j = 0; j < PB; ++j) x[j] = lane < PB ? A[(c0 + lane) + (c0 + j) * LDS] : (T)0;
                const int nk = min(PB, kb - c0);
                // the default pivot rule (no signs, no regularization) without
                // per-pivot branches; signs or a regularization take the general form
                if (F.signs != nullptr || F.reg_kind != 0)
                    pivot_block<T, 0>(x, F, gc0 + c0, nk, chol, cb, pv + c0, rd + c0, lane);
                else if (chol)
                    pivot_block<T, 1>(x, F, gc0 + c0, nk, true, cb, pv + c0, rd + c0, lane);
                else
                    pivot_block<T, 2>(x, F, gc0 + c0, nk, false, cb, pv + c0, rd + c0, lane);
                if (lane < PB)
#pragma unroll
                    for (int j = 0; j < PB; ++j)
                        if (j <= lane) A[(c0 + lane) + (c0 + j) * LDS] = x[j];
            }
            __syncthreads();
            // (2) rows below: y L_D' = a (unit L_D for LDLt), then l = y / d for LDLt
            // (the reciprocals rd loaded unconditionally and applied by selects:
            // a load under a branch would be waited for inside it)
            const int rb = c0 + PB;
            if (tid < 64 && rb + lane < NB) {
                const int i = rb + lane;
                T y[PB], rv[PB];
#pragma unroll
                for (int j = 0; j < PB; ++j) {
                    y[j] = A[i + (c0 + j) * LDS];
                    rv[j] = rd[c0 + j];
                }
#pragma unroll
                for (int k = 0; k < PB; ++k) {
                    y[k] *= chol ? rv[k] : (T)1;
#pragma unroll
                    for (int j = k + 1; j < PB; ++j) y[j] -= y[k] * A[(c0 + j) + (c0 + k) * LDS];
                }
#pragma unroll
                for (int k = 0; k < PB; ++k) A[i + (c0 + k) * LDS] = chol ? y[k] : y[k] * rv[k];
            }
            __syncthreads();
            // (3) trailing update of the lower triangle, A_ij -= sum_k L_ik s_k L_jk,
            // one 16 x 16 block (ti >= tj > b) per wavefront on the matrix cores
            {
                int q = 0;
                for (int ti = b + 1; ti < NB / PB; ++ti)
                    for (int tj = b + 1; tj <= ti; ++tj, ++q) {
                        if ((q & 3) != wave) continue;
                        v4 acc = v4{0, 0, 0, 0};
#pragma unroll
                        for (int s4 = 0; s4 < PB / 4; ++s4) {
                            const int k = 4 * s4 + lk;
                            T av = A[(PB * ti + li) + (c0 + k) * LDS];
                            const T sv = pv[c0 + k];
                            av *= chol ? (T)1 : sv;
                            acc = Mfma<T>::op(av, A[(PB * tj + li) + (c0 + k) * LDS], acc);
                        }
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int row = PB * ti + Mfma<T>::row(lk, v), col = PB * tj + li;
                            if (row >= col) A[row + col * LDS] -= acc[v];
                        }
                    }
            }
            __syncthreads();
        }
        for (int c = cg; c < kb; c += 4) {
            T* colp = blk + (int64_t)c * nj;
            if (r < kb && c <= r) stc(colp + r, A[r + c * LDS]);
        }
        __syncthreads();  // L is read out before the inverse overwrites it
    } else if (tid < NB) {
        rd[tid] = chol ? (T)1 / A[tid + tid * LDS] : (T)1;
    }
    if (invert_only) __syncthreads();
    // W = L^-1 in place.  Diagonal blocks: wavefront w inverts block w; lane
    // c < 16 keeps column c of W_ww (forward substitution, unit for LDLt).
    {
        const int bw = cg, c0 = bw * PB;
        T w[PB];
        if (lane < PB) {
#pragma unroll
            for (int i = 0; i < PB; ++i) w[i] = (i == lane) ? (T)1 : (T)0;
#pragma unroll
            for (int k = 0; k < PB; ++k) {
                const T rk = rd[c0 + k];
                w[k] *= chol ? rk : (T)1;
#pragma unroll
                for (int i = k + 1; i < PB; ++i) w[i] -= A[(c0 + i) + (c0 + k) * LDS] * w[k];
            }
        }
        // a wavefront's LDS reads complete before its writes: only this
        // wavefront touches block (bw, bw) here
        if (lane < PB)
#pragma unroll
            for (int i = 0; i < PB; ++i) A[(c0 + i) + (c0 + lane) * LDS] = (i >= lane) ? w[i] : (T)0;
    }
    __syncthreads();
    for (int bi = 1; bi < NB / PB; ++bi) {
        // T_ij = sum_{k=j}^{i-1} L_ik W_kj for j < i (W_kj rows done, L_ik intact):
        // block bj by wavefront bj
        if (wave < bi) {
            const int bj = wave;
            v4 acc = v4{0, 0, 0, 0};
            blk_mma<T>(A, PB * bi, PB * bj, A, PB * bj, PB * bj, LDS, bi - bj, nullptr, acc);
#pragma unroll
            for (int v = 0; v < 4; ++v) Tm[bj * PB * PB + Mfma<T>::row(lk, v) + li * PB] = acc[v];
        }
        __syncthreads();
        // W_ij = -W_ii T_ij
        if (wave < bi) {
            const int bj = wave;
            v4 acc = v4{0, 0, 0, 0};
            blk_mma<T>(A, PB * bi, PB * bi, Tm + bj * PB * PB, 0, 0, PB, 1, nullptr, acc);
#pragma unroll
            for (int v = 0; v < 4; ++v) A[(PB * bi + Mfma<T>::row(lk, v)) + (PB * bj + li) * LDS] = -acc[v];
        }
        __syncthreads();
    }
    T* W = F.W + f.woff + k0;
    for (int c = cg; c < kb; c += 4) {
        T* colp = W + (int64_t)c * f.nn;
        if (r < kb) stc(colp + r, (c <= r) ? A[r + c * LDS] : (T)0);
    }
}

// Diagonal tile under GMW81 (chol_factor! GMW81 specialization,
// cholesky.jl:739-875): every pivot takes the largest magnitude below it in
// the whole front column, so the task waits for its whole block column (each
// tile final through the previous block column) and reads the rows below the
// tile.  In the tile, right-looking column by column in LDS.  Below it,
// column j as it stands after the tile's pivots 0..j-1 is
//   a_ij - sum_{q<j} a_iq w_q,  w_q = sum_{p=q}^{j-1} l_jp W_pq,
// with a the block column's values (the panel solve has not run yet), l row j
// of the tile's L (unit for LDLt) and W = L^-1, built row by row alongside:
// W_jq = -w_q / l_jj (l_jj = 1 for LDLt), W_jj = 1 / l_jj.  L goes to the
// panel and W to the W pool, as the blocked path writes them.  LDS:
// 2 NB LDS + NB elements and 6 doubles.  Only the GMW81 instance of the
// persistent kernel holds it.
template <typename T>
__device__ __forceinline__ void potrf_gmw(const DevFactor<T>& F, const Front& f, int k0, T* lds,
                                          const int tid) {
    T* A = lds;                // the tile, 64 x LDS, lower triangle
    T* Wt = lds + NB * LDS;    // W = L^-1 of the tile, row by row
    T* w = Wt + NB * LDS;      // w of the current pivot
    double* s_red = (double*)(w + NB);  // per-wave maxima, then the pivot's l and 1 / l
    double& s_l = s_red[4];
    double& s_inv = s_red[5];
    const int nj = f.nn + f.na, kb = min(NB, f.nn - k0), gc0 = f.c0 + k0, nbelow = nj - k0;
    T* blk = F.panel + f.poff + k0 + (int64_t)k0 * nj;
    const int r = tid & 63, cg = tid >> 6, lane = tid & 63, wave = tid >> 6;
    const bool chol = F.diag == 0;
    for (int c = cg; c < NB; c += 4) {
        A[r + c * LDS] = (r < kb && c < kb && c <= r) ? ldc(blk + r + (int64_t)c * nj) : (r == c ? (T)1 : (T)0);
        Wt[r + c * LDS] = (T)0;
    }
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
        // (1) w_q = sum_{p=q}^{j-1} l_jp W_pq, one thread per q < j
        if (tid < j) {
            T acc = (T)0;
            for (int p = tid; p < j; ++p) acc += A[j + p * LDS] * Wt[p + tid * LDS];
            w[tid] = acc;
        }
        __syncthreads();
        // (2) the largest magnitude below the pivot: tile rows from LDS, rows
        // below the tile from the block column
        double m = (r > j && r < kb && cg == 0) ? fabs((double)A[r + j * LDS]) : 0.0;
        // RB rows per thread and QB columns per step: RB * QB independent
        // loads in flight (the rows are device-coherent loads from far
        // memory); out-of-range rows and columns read row kb / column 0 and
        // are dropped
        constexpr int RB = 4, QB = 4;
        for (int i0 = kb; i0 < nbelow; i0 += 256 * RB) {
            T acc[RB];
            int64_t ro[RB];
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int i = i0 + tid + 256 * u;
                ro[u] = i < nbelow ? i : kb;
                acc[u] = ldc(blk + ro[u] + (int64_t)j * nj);
            }
            for (int q0 = 0; q0 < j; q0 += QB) {
                T a[QB][RB], wq[QB];
#pragma unroll
                for (int qq = 0; qq < QB; ++qq) {
                    const int q = q0 + qq < j ? q0 + qq : 0;
                    wq[qq] = q0 + qq < j ? w[q] : (T)0;
#pragma unroll
                    for (int u = 0; u < RB; ++u) a[qq][u] = ldc(blk + ro[u] + (int64_t)q * nj);
                }
#pragma unroll
                for (int qq = 0; qq < QB; ++qq)
#pragma unroll
                    for (int u = 0; u < RB; ++u) acc[u] -= a[qq][u] * wq[qq];
            }
#pragma unroll
            for (int u = 0; u < RB; ++u)
                if (i0 + tid + 256 * u < nbelow) m = jmax(m, fabs((double)acc[u]));
        }
        m = wave_max(m);
        if (lane == 0) s_red[wave] = m;
        __syncthreads();
        if (tid == 0) {
            const double amax = jmax(jmax(s_red[0], s_red[1]), jmax(s_red[2], s_red[3]));
            const double piv = regularize_gmw(F, sign_of(F, gc0 + j), (double)A[j + j * LDS], amax);
            if (chol) {
                const double l = sqrt(piv);
                s_l = l;
                s_inv = 1.0 / l;
            } else {
                stc(F.d + gc0 + j, (T)piv);
                s_l = piv;
                s_inv = 1.0 / piv;
            }
        }
        __syncthreads();
        // (3) column j of L (the LDLt diagonal keeps the unregularized value)
        // and row j of W
        const T inv = (T)s_inv, pj = (T)s_l;
        if (cg == 0) {
            if (r > j && r < kb) A[r + j * LDS] *= inv;
            if (r == j && chol) A[j + j * LDS] = pj;
        } else if (cg == 1 && r <= j) {
            Wt[j + r * LDS] = r == j ? (chol ? inv : (T)1) : -w[r] * (chol ? inv : (T)1);
        }
        __syncthreads();
        // (4) rank-1 update of the trailing tile: a_rc -= l_rj (d_j) l_cj
        if (r > j && r < kb) {
            const T lr = A[r + j * LDS] * (chol ? (T)1 : pj);
            for (int c = j + 1 + cg; c <= r; c += 4) A[r + c * LDS] -= lr * A[c + j * LDS];
        }
        __syncthreads();
    }
    for (int c = cg; c < kb; c += 4)
        if (r < kb && c <= r) stc(blk + r + (int64_t)c * nj, A[r + c * LDS]);
    T* W = F.W + f.woff + k0;
    for (int c = cg; c < kb; c += 4)
        if (r < kb) stc(W + r + (int64_t)c * f.nn, (c <= r) ? Wt[r + c * LDS] : (T)0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_diag(DevPlan P, DevFactor<T> F, int t0, int invert_only) {
    __shared__ T lds[PTILE];
    const DiagTask dt = P.diag[t0 + blockIdx.x];
    potrf_blocked<T>(F, P.fronts[dt.front], dt.k0, invert_only != 0, lds, (int)threadIdx.x);
}

template <typename T>
__device__ __forceinline__ T* arr_base(const DevFactor<T>& F, int a) {
    return a == ARR_PANEL ? F.panel : (a == ARR_U ? F.U : F.W);
}

// One 64x64 output tile; wave w owns a 32x32 quarter as 2x2 blocks of 16x16
// MFMA tiles (v_mfma_*_16x16x4: lane l holds A[l%16][l/16] and B[l/16][l%16];
// the row of output register v is Mfma<T>::row).  The k-loop is software
// pipelined: the next KC-chunk of A and B is loaded into registers while the
// MFMAs consume the current chunk from LDS (As/Bs: KC x LDT each).  For
// C -= A B' the C tile is fetched before the k-loop.
template <typename T>
__device__ __forceinline__ void gemm_tile(const DevFactor<T>& F, const GemmTask& g, T* As, T* Bs, const int tid) {
    typedef typename Mfma<T>::v4 v4;
    constexpr int PER = KC * 64 / 256;  // elements of A (and of B) per thread and chunk
    const T* A = arr_base(F, g.arrA) + g.a_off;
    const T* B = arr_base(F, g.arrB) + g.b_off;
    T* C = arr_base(F, g.arrC) + g.c_off;
    const int wave = tid >> 6, lane = tid & 63;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    const bool sub = g.mode == G_SUB;
    const T* s = (F.diag != 0 && sub && g.s_col >= 0) ? F.d + g.s_col : nullptr;
    T cv[2][2][4];
    if (sub) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = wr + i * 16 + Mfma<T>::row(lk, v), col = wc + j * 16 + li;
                    cv[i][j][v] = (row < g.m && col < g.n && (!g.lower || row >= col))
                                      ? ldc(C + row + (int64_t)col * g.ldc) : (T)0;
                }
    }
    v4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4{0, 0, 0, 0};
    T ra[PER], rb[PER], rs[PER];
    const int r = tid & 63, kq = tid >> 6;  // element (r, kq + 4i) of each chunk
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int gk = k0 + kq + 4 * i;
            T av = (T)0, bv = (T)0, sv = (T)1;
            if (gk < g.k) {
                if (r < g.m) av = ldc(A + r + (int64_t)gk * g.lda);
                if (s) sv = ldc(s + gk);  // applied when staged: no wait here
                if (r < g.n) bv = ldc(B + r + (int64_t)gk * g.ldb);
            }
            ra[i] = av;
            rb[i] = bv;
            rs[i] = sv;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < g.k; k0 += KC) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            As[(kq + 4 * i) * LDT + r] = s ? ra[i] * rs[i] : ra[i];
            Bs[(kq + 4 * i) * LDT + r] = rb[i];
        }
        __syncthreads();
        if (k0 + KC < g.k) fetch(k0 + KC);
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
    // the LDLt pivots of this thread's two columns, loaded before the stores
    // (unconditionally, column clamped) rather than one wait per element
    T dcol[2] = {(T)1, (T)1};
    if (scale)
#pragma unroll
        for (int j = 0; j < 2; ++j) dcol[j] = ldc(F.d + g.s_col + min(wc + j * 16 + li, g.n - 1));
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
                    if (sub) {
                        stc(cp, cv[i][j][v] - acc[i][j][v]);
                    } else {
                        T x = acc[i][j][v];
                        if (scale) x /= dcol[j];
                        stc(cp, x);
                    }
                }
            }
        }
    }
}

constexpr int KW = 16;        // k-chunk of the wide tile
constexpr int LDW = 128 + 4;  // leading dimension of its LDS stages

// 128 x 128 update-matrix tile, C -= A diag(s) B' (lower part only on the
// diagonal): twice the flops per byte of the 64 x 64 tile, for the long-k
// update matrices that carry most of the flops.  Wave w owns the 64 x 64
// quarter (w & 1, w >> 1) as 4 x 4 MFMA blocks (the strictly upper quarter
// of a diagonal tile does no MFMA work).  The k-loop stages KW columns of A
// and B through two LDS buffers (As/Bs: 2 x KW x LDW each): while the waves
// multiply out of one buffer, the next chunk, prefetched into registers one
// chunk ahead, is written to the other, so each chunk costs one barrier.  C
// is updated after the k-loop in four row-block rounds.
template <typename T>
__device__ __forceinline__ void gemm_tile128(const DevFactor<T>& F, const GemmTask& g, T* As, T* Bs, const int tid) {
    typedef typename Mfma<T>::v4 v4;
    constexpr int PER = KW * 128 / 256;
    const T* A = arr_base(F, g.arrA) + g.a_off;
    const T* B = arr_base(F, g.arrB) + g.b_off;
    T* C = arr_base(F, g.arrC) + g.c_off;
    const int wave = tid >> 6, lane = tid & 63;
    const int wr = (wave & 1) * 64, wc = (wave >> 1) * 64;
    const int li = lane & 15, lk = lane >> 4;
    const bool idle = g.lower && wc > wr;  // strictly above the diagonal
    const T* s = (F.diag != 0 && g.s_col >= 0) ? F.d + g.s_col : nullptr;
    v4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4{0, 0, 0, 0};
    T ra[PER], rb[PER], rs[PER];
    const int r = tid & 127, kq = tid >> 7;  // element (r, kq + 2i) of each chunk
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int gk = k0 + kq + 2 * i;
            T av = (T)0, bv = (T)0, sv = (T)1;
            if (gk < g.k) {
                if (r < g.m) av = ldc(A + r + (int64_t)gk * g.lda);
                if (s) sv = ldc(s + gk);  // applied when staged: no wait here
                if (r < g.n) bv = ldc(B + r + (int64_t)gk * g.ldb);
            }
            ra[i] = av;
            rb[i] = bv;
            rs[i] = sv;
        }
    };
    auto stage = [&](int buf) {
        T* a = As + buf * KW * LDW;
        T* b = Bs + buf * KW * LDW;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            a[(kq + 2 * i) * LDW + r] = s ? ra[i] * rs[i] : ra[i];
            b[(kq + 2 * i) * LDW + r] = rb[i];
        }
    };
    fetch(0);
    stage(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < g.k; k0 += KW, buf ^= 1) {
        const bool more = k0 + KW < g.k;
        if (more) fetch(k0 + KW);
        if (!idle) {
            const T* a0 = As + buf * KW * LDW;
            const T* b0 = Bs + buf * KW * LDW;
#pragma unroll
            for (int kk = 0; kk < KW; kk += 4) {
                T a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = a0[(kk + lk) * LDW + wr + 16 * i + li];
                    b[i] = b0[(kk + lk) * LDW + wc + 16 * i + li];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(a[i], b[j], acc[i][j]);
            }
        }
        // the other buffer was last read in the previous chunk, before its barrier
        if (more) stage(buf ^ 1);
        __syncthreads();
    }
    if (idle) return;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        T cv[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wr + 16 * i + Mfma<T>::row(lk, v), col = wc + 16 * j + li;
                cv[j][v] = (row < g.m && col < g.n && (!g.lower || row >= col))
                               ? ldc(C + row + (int64_t)col * g.ldc) : (T)0;
            }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wr + 16 * i + Mfma<T>::row(lk, v), col = wc + 16 * j + li;
                if (row < g.m && col < g.n && (!g.lower || row >= col))
                    stc(C + row + (int64_t)col * g.ldc, cv[j][v] - acc[i][j][v]);
            }
    }
}

// Dependency q of DAG task d: the counter w that must reach target.  Returns
// false when d has no dependency q; *nw is the number of dependencies.
template <typename T, bool GMW>
__device__ __forceinline__ bool dag_dep(const DevPlan& P, const DevFactor<T>& F, const DevDag& D, const DagTask& d,
                        int q, int32_t** w, int* target, int* nw) {
    const int J = d.front;
    const DagFront g = D.df[J];
    int32_t* cnt = D.cnt + g.toff;
    switch (d.kind) {
        case DT_EA:
        case DT_SMALL: {  // children complete, earlier users of the U slot consumed
            const Front f = P.fronts[J];
            const int nchd = f.chd1 - f.chd0, u0 = D.uwait_ptr[J];
            *nw = nchd + D.uwait_ptr[J + 1] - u0;
            if (q >= *nw) return false;
            if (q < nchd) {
                const int K = P.chd[f.chd0 + q];
                *w = F.flags + K;
                *target = D.df[K].n_done;
            } else {
                const int c = D.uwait[u0 + q - nchd];
                *w = D.eacnt + c;
                *target = D.df[c].n_ea;
            }
            return true;
        }
        case DT_POTRF:
            // the diagonal tile (the fused products wait for their own tiles
            // inside the task, after the factorization); under GMW81 every
            // tile of the block column, updated through the previous column
            *nw = (GMW && d.a > 0) ? g.R - d.a : 1;
            if (q >= *nw) return false;
            *w = d.a == 0 ? D.eapcnt + J : cnt + d.a * g.R + d.a + q;
            *target = d.a == 0 ? g.n_eap : d.a;
            return true;
        case DT_TRSM:
            *nw = d.c > 0 ? 2 : 1;
            if (q >= *nw) return false;
            *w = cnt + d.c * g.R + (q == 0 ? d.c : d.b);
            *target = q == 0 ? d.c + 1 : d.c;
            return true;
        case DT_UPD: {  // columns [K0, K1): rows I, Jc final through K1 - 1
            const int K0 = d.d & 0xffff, K1 = d.d >> 16;
            *nw = K0 > 0 ? 3 : 2;
            if (q >= *nw) return false;
            *w = q == 2 ? cnt + d.c * g.R + d.b : cnt + (K1 - 1) * g.R + (q == 0 ? d.b : d.c);
            *target = q == 2 ? K0 : K1;
            return true;
        }
        default: {  // DT_SYRK: the tile's row blocks of the panel are final, and
                    // every extend-add of the front (the update-matrix columns) is done
            *nw = 3 + (d.d & 1) + ((d.d >> 1) & 1);
            if (q >= *nw) return false;
            if (q == *nw - 1) {
                *w = D.eacnt + J;
                *target = g.n_ea;
                return true;
            }
            const int blk = q == 0 ? d.b : q == 1 ? d.c : (q == 2 && (d.d & 1)) ? d.b + 1 : d.c + 1;
            *w = cnt + (g.p - 1) * g.R + blk;
            *target = g.p;
            return true;
        }
    }
}

// Tickets a workgroup holds at once, and lanes of wave 0 per held ticket
// (one dependency each).  Holding more than one ticket lets a workgroup run a
// later ready ticket first, but every held ticket then waits for its holder:
// measured on the 64^3 7-point Laplacian the launch took 36.4 ms with one
// ticket, 39.2 ms with two, 41.5 ms with four and 44.3 ms with eight.
constexpr int WIN = 1;
constexpr int WDEP = 64;
constexpr unsigned long long WMASK = WDEP == 64 ? ~0ull : (1ull << (WDEP & 63)) - 1;

// Persistent tile-DAG factorization of the big region.  Tickets are handed
// out in the planner's order (a list schedule, critical path first); a task
// depends only on tasks with smaller tickets.  Each workgroup holds WIN
// tickets and runs whichever of them is ready (the oldest first), so it
// blocks only when none of its tickets is ready.  The smallest unfinished
// ticket always belongs to some workgroup and is ready (its dependencies all
// have smaller tickets), so the launch is deadlock-free on any number of
// resident workgroups.  Counters:
//   F.flags[J]   finished update-matrix tiles of front J (1 for a small front)
//   D.eacnt[J]   finished extend-add tasks of J (1 for a small front)
//   D.eapcnt[J]  finished extend-add tasks of J into panel columns (the
//                first diagonal tile waits only for these; the update-matrix
//                tiles wait for all of them)
//   D.cnt[toff_J + K*R_J + I]  panel columns applied to tile (I, K); K+1
//                once the tile is final
// Synchronization per task: lanes of wave 0 poll the dependencies of the
// held tickets with relaxed loads (WDEP lanes per ticket).  Every value a
// task reads that another task of this launch wrote is loaded with a
// device-coherent (sc1) load, and every task output is stored with a
// device-coherent store, so no L2 write-back or invalidate is needed: at the
// end every wave drains its stores, the workgroup meets at a barrier, and
// thread 0 publishes with relaxed counter updates.  (Whole-L2 write-backs per
// task cost 3 ms of the 35 ms launch on the 64^3 Laplacian.)
// GMW: the GMW81 instance (diagonal tiles by potrf_gmw, block-column waits);
// the other regularizations run the instance without it.
template <typename T, bool GMW>
__global__ __launch_bounds__(256, 2) void k_dag(DevPlan P, DevFactor<T> F, DevDag D) {
    constexpr int S1 = PTILE, S2 = 2 * KC * LDT, S3 = 4 * KW * LDW;
    constexpr int SMEM = S1 > S2 ? (S1 > S3 ? S1 : S3) : (S2 > S3 ? S2 : S3);
    static_assert(SMEM >= 2 * NB * LDS + NB + 6 * (int)(sizeof(double) / sizeof(T)), "potrf_gmw LDS");
    __shared__ __align__(16) T smem[SMEM];
    __shared__ T s_piv, s_inv;
    __shared__ int s_fail, s_sel;
    __shared__ int s_win[WIN];
    if (threadIdx.x == 0)
        for (int i = 0; i < WIN; ++i) {
            const int t = (int)atomicAdd(F.ticket, 1u);
            s_win[i] = t < D.ntask ? t : -1;
        }
    __syncthreads();
    // wave 0, lane (slot, q): dependency q of the ticket held in slot
    int my_t = -3, my_tgt = 0, my_extra = 0;
    bool my_has = false;
    int32_t* my_w = nullptr;
    unsigned long long t_mark = 0;
    if (D.prof && threadIdx.x == 0) t_mark = wall_clock64();
    for (;;) {
        // an opaque copy of the thread index: values derived from it are
        // recomputed per task instead of being held (and spilled) across the
        // whole persistent loop
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        if (tid < 64) {
            const int slot = tid / WDEP, q = tid & (WDEP - 1);
            int sel = -1;
            Watch watch;
            for (int spins = 0;; ++spins) {
                const int t = s_win[slot];
                if (t != my_t) {  // a new ticket in this slot
                    my_t = t;
                    my_has = false;
                    my_extra = 0;
                    if (t >= 0) {
                        const DagTask dd = D.task[t];
                        int nw = 0;
                        my_has = dag_dep<T, GMW>(P, F, D, dd, q, &my_w, &my_tgt, &nw);
                        if (q == 0 && nw > WDEP) my_extra = nw;
                    }
                }
                bool ok = !my_has || ld_rel(my_w) >= my_tgt;
                if (my_extra) {  // more dependencies than lanes (many children): checked here
                    const DagTask dd = D.task[my_t];
                    for (int k = WDEP; k < my_extra && ok; ++k) {
                        int32_t* w2;
                        int tg2, nw2;
                        if (dag_dep<T, GMW>(P, F, D, dd, k, &w2, &tg2, &nw2)) ok = ld_rel(w2) >= tg2;
                    }
                }
                const unsigned long long okm = __ballot(ok);
                if (tid == 0) {
                    int best = -1, bt = 0x7fffffff;
                    bool any = false;
                    for (int i = 0; i < WIN; ++i) {
                        const int ti = s_win[i];
                        if (ti < 0) continue;
                        any = true;
                        if (((okm >> (WDEP * i)) & WMASK) == WMASK && ti < bt) {
                            bt = ti;
                            best = i;
                        }
                    }
                    sel = any ? best : -2;
                    // the watchdog of mf_sync.h: progress is a satisfied
                    // dependency more or a ticket taken anywhere in the launch
                    if (sel == -1 && (spins & 255) == 255) {
                        if (ld_rel(F.hang)) {
                            sel = -2;
                        } else if (watch.expired(__popcll(okm), F.ticket, F.stall)) {
                            give_up(F.hang, my_w, my_tgt, my_w ? ld_rel(my_w) : 0, my_t);
                            sel = -2;
                        }
                    }
                }
                sel = __shfl(sel, 0);
                if (sel != -1) break;
                __builtin_amdgcn_s_sleep(1);
            }
            if (tid == 0) {
                s_sel = sel;
                s_fail = 0;
                const unsigned long long a0 = D.prof ? wall_clock64() : 0;
                // no acquire: every read of data this launch wrote is a
                // device-coherent load (ldc), so no cache needs an invalidate
                if (D.prof) atomicAdd(D.prof + 3 * DT_KINDS + 1, wall_clock64() - a0);
            }
        }
        __syncthreads();
        const int sel = s_sel;
        if (sel < 0) break;
        const int t = s_win[sel];
        const DagTask d = D.task[t];
        const int J = d.front;
        const DagFront g = D.df[J];
        int32_t* cnt = D.cnt + g.toff;
        unsigned long long t_begin = 0;
        if (D.prof && tid == 0) t_begin = wall_clock64();
        // critical-chain tasks may issue ahead of the co-resident workgroup's
        // matrix-core waves (wave priority; uniform branch)
        const bool raise = ((D.prio >> d.kind) & 1) ||
                           (((D.prio >> DT_KINDS) & 1) && d.kind == DT_UPD && d.c == (d.d >> 16));
        if (raise) __builtin_amdgcn_s_setprio(3);
        switch (d.kind) {
            case DT_EA: {
                const int wave = tid >> 6, lane = tid & 63;
                if (d.a + wave < d.b) ea_columns<T>(P, F, d.a + wave, d.b, smem + wave * EA_WIN, lane);
                break;
            }
            case DT_SMALL:
                small_front<T>(P, F, J, smem, &s_piv, &s_inv, &s_fail, tid);
                break;
            case DT_POTRF:
                if constexpr (GMW) potrf_gmw<T>(F, P.fronts[J], d.a * NB, smem, tid);
                else potrf_blocked<T>(F, P.fronts[J], d.a * NB, false, smem, tid);
                if (d.b >= 0) {
                    // publish the diagonal tile (the other panel solves of column a
                    // need only it), then the fused panel solve of tile (a+1, a)
                    drain_stores();
                    __syncthreads();
                    if (tid == 0) {
                        __builtin_amdgcn_s_waitcnt(0);
                        __hip_atomic_store(cnt + d.a * g.R + d.a, d.a + 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                        // tile (a+1, a) updated through column a - 1: set by a
                        // task with a smaller ticket (the planner checks), so
                        // the wait ends
                        (void)bounded_wait(cnt + d.a * g.R + d.a + 1, d.a, F.hang, F.ticket, F.stall, 1);
                    }
                    __syncthreads();
                    gemm_tile<T>(F, P.gemm[d.b], smem, smem + KC * LDT, tid);
                    if (d.c >= 0) {
                        // publish tile (a+1, a), then the fused look-ahead update of
                        // the next diagonal tile, which reads it back
                        drain_stores();
                        __syncthreads();
                        if (tid == 0) {
                            __builtin_amdgcn_s_waitcnt(0);
                            __hip_atomic_store(cnt + d.a * g.R + d.a + 1, d.a + 1, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                            // tile (a+1, a+1) updated through column a - 1 (as above)
                            (void)bounded_wait(cnt + (d.a + 1) * g.R + d.a + 1, d.a, F.hang, F.ticket,
                                               F.stall, 1);
                        }
                        __syncthreads();
                        gemm_tile<T>(F, P.gemm[d.c], smem, smem + KC * LDT, tid);
                    }
                }
                break;
            default: {
                const GemmTask gt = P.gemm[d.a];
                if (gt.wide) gemm_tile128<T>(F, gt, smem, smem + 2 * KW * LDW, tid);
                else gemm_tile<T>(F, gt, smem, smem + KC * LDT, tid);
                break;
            }
        }
        if (raise) __builtin_amdgcn_s_setprio(0);
        drain_stores();
        __syncthreads();
        if (tid == 0) {
            if (D.prof) {
                const unsigned long long now = wall_clock64();
                if (D.trace) {
                    D.trace[2 * t] = t_begin | ((unsigned long long)blockIdx.x << 48);
                    D.trace[2 * t + 1] = now;
                }
                atomicAdd(D.prof + 3 * d.kind, t_begin - t_mark);
                atomicAdd(D.prof + 3 * d.kind + 1, now - t_begin);
                atomicAdd(D.prof + 3 * d.kind + 2, 1ull);
                t_mark = now;
            }
            const unsigned long long r0 = D.prof ? wall_clock64() : 0;
            // every task output is written through (device-coherent stores), and
            // every wave drained its stores before the barrier: no L2 write-back
            __builtin_amdgcn_s_waitcnt(0);
            constexpr int R = __ATOMIC_RELAXED, S = __HIP_MEMORY_SCOPE_AGENT;
            switch (d.kind) {
                case DT_EA:
                    if (d.c) __hip_atomic_fetch_add(D.eapcnt + J, 1, R, S);
                    __hip_atomic_fetch_add(D.eacnt + J, 1, R, S);
                    break;
                case DT_SMALL:
                    __hip_atomic_store(D.eacnt + J, 1, R, S);
                    __hip_atomic_store(F.flags + J, 1, R, S);
                    break;
                case DT_POTRF:
                    __hip_atomic_store(cnt + d.a * g.R + d.a, d.a + 1, R, S);
                    if (d.b >= 0) __hip_atomic_store(cnt + d.a * g.R + d.a + 1, d.a + 1, R, S);
                    if (d.c >= 0) __hip_atomic_store(cnt + (d.a + 1) * g.R + d.a + 1, d.a + 1, R, S);
                    break;
                case DT_TRSM: __hip_atomic_store(cnt + d.c * g.R + d.b, d.c + 1, R, S); break;
                case DT_UPD: __hip_atomic_store(cnt + d.c * g.R + d.b, d.d >> 16, R, S); break;
                default: __hip_atomic_fetch_add(F.flags + J, 1, R, S); break;
            }
            const unsigned long long r1 = D.prof ? wall_clock64() : 0;
            const int nt = (int)atomicAdd(F.ticket, 1u);
            s_win[sel] = nt < D.ntask ? nt : -1;
            if (D.prof) {
                atomicAdd(D.prof + 3 * DT_KINDS, r1 - r0);
                atomicAdd(D.prof + 3 * DT_KINDS + 2, wall_clock64() - r1);
            }
        }
        // closes the thread-0 block before the back-edge
        __syncthreads();
    }
}

template <typename T>
__global__ void k_fill(T* p, int64_t count, T value) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = value;
}

// Reference-layout Dval/Lval (assembled A, or a factor) -> internal panels
// (zeroed before); one workgroup per global column.  uplo 0 (:L): the
// diagonal block is nn x nn lower and the off-diagonal block na x nn;
// uplo 1 (:U): both transposed, the block above being nn x na
// (chordal_triangular.jl:852-926, 1015-1055).  adiag[g] receives the
// diagonal entry of column g.
template <typename T>
__global__ __launch_bounds__(256) void k_assemble_ref(DevPlan P, DevFactor<T> F,
                                                      const double* Dval, const double* Lval,
                                                      int uplo, double* adiag) {
    const int g = blockIdx.x;
    const OutFront o = P.out[P.col2ref[g]];
    const Front f = P.fronts[o.J];
    const int nj = f.nn + f.na, c = g - o.r0, cl = o.r0 - f.c0;
    T* col = F.panel + f.poff + (int64_t)(cl + c) * nj;
    if (uplo == 0) {
        for (int r = c + threadIdx.x; r < o.nn; r += blockDim.x)
            col[cl + r] = (T)Dval[o.Dp + r + (int64_t)c * o.nn];
        for (int i = threadIdx.x; i < o.na; i += blockDim.x)
            col[P.septpos[o.sepoff + i]] = (T)Lval[o.Lp + i + (int64_t)c * o.na];
    } else {
        for (int r = c + threadIdx.x; r < o.nn; r += blockDim.x)
            col[cl + r] = (T)Dval[o.Dp + c + (int64_t)r * o.nn];
        for (int i = threadIdx.x; i < o.na; i += blockDim.x)
            col[P.septpos[o.sepoff + i]] = (T)Lval[o.Lp + c + (int64_t)i * o.nn];
    }
    if (threadIdx.x == 0) adiag[g] = Dval[o.Dp + c + (int64_t)c * o.nn];
}

// setflatindex! for every nonzero into the panels (tp: panel slot or -1);
// td: the permuted column of a diagonal entry or -1 (adiag keeps the
// assembled diagonal for the LDLt copy-out of nn = 1 fronts).
template <typename T>
__global__ void k_refresh(DevFactor<T> F, const int64_t* tp, const int64_t* td, const double* x,
                          int64_t nnz, double* adiag) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const double v = x[k];
    const int64_t p = tp[k];
    if (p >= 0) F.panel[p] = (T)v;
    const int64_t q = td[k];
    if (q >= 0) adiag[q] = v;
}

// Internal panels -> the reference's Dval/Lval/d; one workgroup per global
// column.  Every entry of the column's blocks is written (zeros outside the
// stored triangle), so the arrays need no clearing before.  uplo as for
// k_assemble_ref.
template <typename T>
__global__ __launch_bounds__(256) void k_copyout(DevPlan P, DevFactor<T> F, double* Dval,
                                                 double* Lval, double* dout, int uplo,
                                                 const double* adiag) {
    const int g = blockIdx.x;
    const OutFront o = P.out[P.col2ref[g]];
    const Front f = P.fronts[o.J];
    const int nj = f.nn + f.na, c = g - o.r0, cl = o.r0 - f.c0;
    const T* col = F.panel + f.poff + (int64_t)(cl + c) * nj;
    const bool keep_diag = (F.diag != 0 && o.nn == 1);  // chol_loop_nod!: Dval keeps A's entry
    if (Dval) {
        for (int r = threadIdx.x; r < o.nn; r += blockDim.x) {
            double v = r >= c ? (double)col[cl + r] : 0.0;
            if (r == c && keep_diag) v = adiag[g];
            if (uplo == 0) Dval[o.Dp + r + (int64_t)c * o.nn] = v;
            else Dval[o.Dp + c + (int64_t)r * o.nn] = v;
        }
    }
    if (Lval) {
        for (int i = threadIdx.x; i < o.na; i += blockDim.x) {
            const double v = (double)col[P.septpos[o.sepoff + i]];
            if (uplo == 0) Lval[o.Lp + i + (int64_t)c * o.na] = v;
            else Lval[o.Lp + c + (int64_t)i * o.nn] = v;
        }
    }
    if (dout && F.diag != 0 && threadIdx.x == 0) dout[g] = (double)F.d[g];
}

// dst[i] = (double)src[i]: the FP64 copy of an FP32 factor's panels that
// the FP64-arithmetic solves read.
__global__ void k_widen(const float* src, double* dst, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) dst[i] = (double)src[i];
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
                         const double* Lval, int uplo, double* adiag, hipStream_t s) {
    if (P.n <= 0) return;
    hipLaunchKernelGGL(k_assemble_ref<T>, dim3(P.n), dim3(256), 0, s, P, F, Dval, Lval, uplo,
                       adiag);
}

template <typename T>
void launch_refresh(const DevFactor<T>& F, const int64_t* tp, const int64_t* td,
                    const double* nzval, int64_t nnz, double* adiag, hipStream_t s) {
    if (nnz <= 0) return;
    hipLaunchKernelGGL(k_refresh<T>, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s, F, tp,
                       td, nzval, nnz, adiag);
}

template <typename T>
void launch_small(const DevPlan& P, const DevFactor<T>& F, const int32_t* list, const int32_t* units,
                  int count, hipStream_t s) {
    if (count <= 0) return;
    const int grid = count < 2048 ? count : 2048;
    hipLaunchKernelGGL(k_small<T>, dim3(grid), dim3(256), 0, s, P, F, list, units, count);
}

template <typename T>
void launch_diag(const DevPlan& P, const DevFactor<T>& F, int t0, int t1, bool invert_only,
                 hipStream_t s) {
    if (t1 <= t0) return;
    hipLaunchKernelGGL(k_diag<T>, dim3(t1 - t0), dim3(256), 0, s, P, F, t0, invert_only ? 1 : 0);
}

template <typename T>
void launch_dag(const DevPlan& P, const DevFactor<T>& F, const DevDag& D, hipStream_t s) {
    if (D.ntask <= 0) return;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const bool gmw = F.reg_kind == 2;
    if (gmw) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_dag<T, true>, 256, 0);
    else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_dag<T, false>, 256, 0);
    if (D.wg_per_cu > 0 && D.wg_per_cu < per) per = D.wg_per_cu;
    int grid = std::max(1, cus) * std::max(1, per);
    if (grid > D.ntask) grid = D.ntask;
    if (gmw) hipLaunchKernelGGL((k_dag<T, true>), dim3(grid), dim3(256), 0, s, P, F, D);
    else hipLaunchKernelGGL((k_dag<T, false>), dim3(grid), dim3(256), 0, s, P, F, D);
}

// gmw81_beta (gmw81.jl:74-101): the largest diagonal and off-diagonal
// magnitudes of the assembled matrix (the panels hold A before the
// factorization), one workgroup per internal front, combined with atomicMax
// on the bit patterns (non-negative doubles order as their bits; a NaN
// orders above every number and propagates).
template <typename T>
__global__ __launch_bounds__(256) void k_gmw_reduce(DevPlan P, DevFactor<T> F, unsigned long long* acc) {
    const Front f = P.fronts[blockIdx.x];
    const int nj = f.nn + f.na;
    const T* pn = F.panel + f.poff;
    double g = 0.0, x = 0.0;
    const int64_t total = (int64_t)nj * f.nn;
    for (int64_t e = threadIdx.x; e < total; e += 256) {
        const int c = (int)(e / nj), r = (int)(e - (int64_t)c * nj);
        if (r < c) continue;
        const double v = fabs((double)pn[e]);
        if (r == c) g = jmax(g, v);
        else x = jmax(x, v);
    }
    g = wave_max(g);
    x = wave_max(x);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(acc, (unsigned long long)__builtin_bit_cast(unsigned long long, g));
        atomicMax(acc + 1, (unsigned long long)__builtin_bit_cast(unsigned long long, x));
    }
}

// beta = max(gamma, xi / sqrt(n^2 - 1), eps(T)), rounded to T
__global__ void k_gmw_final(double* gmw, double n, double eps_t, int f32) {
    const unsigned long long* acc = (const unsigned long long*)gmw;
    const double gamma = __builtin_bit_cast(double, acc[0]), xi = __builtin_bit_cast(double, acc[1]);
    const double b = jmax(jmax(gamma, xi / sqrt(n * n - 1.0)), eps_t);
    gmw[2] = f32 ? (double)(float)b : b;
}

template <typename T>
void launch_gmw_beta(const DevPlan& P, const DevFactor<T>& F, double* gmw, int64_t n, hipStream_t s) {
    launch_fill<unsigned long long>((unsigned long long*)gmw, 2, 0ull, s);
    if (P.nJ > 0)
        hipLaunchKernelGGL(k_gmw_reduce<T>, dim3(P.nJ), dim3(256), 0, s, P, F, (unsigned long long*)gmw);
    const double eps_t = sizeof(T) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
    hipLaunchKernelGGL(k_gmw_final, dim3(1), dim3(1), 0, s, gmw, (double)n, eps_t, sizeof(T) == 4 ? 1 : 0);
}

void launch_widen(const float* src, double* dst, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_widen, dim3(grid_for(count, 256)), dim3(256), 0, s, src, dst, count);
}

template <typename T>
void launch_copyout(const DevPlan& P, const DevFactor<T>& F, double* Dval, double* Lval,
                    double* dout, int uplo, const double* adiag, hipStream_t s) {
    if (P.n <= 0) return;
    hipLaunchKernelGGL(k_copyout<T>, dim3(P.n), dim3(256), 0, s, P, F, Dval, Lval, dout, uplo,
                       adiag);
}

#define MF_INST(T)                                                                              \
    template void launch_fill<T>(T*, int64_t, T, hipStream_t);                                  \
    template void launch_assemble_ref<T>(const DevPlan&, const DevFactor<T>&, const double*,    \
                                         const double*, int, double*, hipStream_t);             \
    template void launch_refresh<T>(const DevFactor<T>&, const int64_t*, const int64_t*,        \
                                    const double*, int64_t, double*, hipStream_t);              \
    template void launch_small<T>(const DevPlan&, const DevFactor<T>&, const int32_t*,          \
                                  const int32_t*, int, hipStream_t);                            \
    template void launch_diag<T>(const DevPlan&, const DevFactor<T>&, int, int, bool,           \
                                 hipStream_t);                                                 \
    template void launch_copyout<T>(const DevPlan&, const DevFactor<T>&, double*, double*,      \
                                    double*, int, const double*, hipStream_t);                  \
    template void launch_dag<T>(const DevPlan&, const DevFactor<T>&, const DevDag&, hipStream_t); \
    template void launch_gmw_beta<T>(const DevPlan&, const DevFactor<T>&, double*, int64_t, hipStream_t);
MF_INST(double)
MF_INST(float)
template void launch_fill<int32_t>(int32_t*, int64_t, int32_t, hipStream_t);
template void launch_fill<uint32_t>(uint32_t*, int64_t, uint32_t, hipStream_t);
template void launch_fill<unsigned long long>(unsigned long long*, int64_t, unsigned long long,
                                              hipStream_t);

}  // namespace mf
