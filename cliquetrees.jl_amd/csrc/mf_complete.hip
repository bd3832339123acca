// mf_complete.hip — gfx950 kernels of the completion (complete_impl!,
// src/Multifrontal.jl/src/complete.jl:46-451): Dval/Lval hold a matrix Sigma
// on the factor's pattern; the result is the Cholesky factor of the inverse
// of Sigma's maximum-determinant positive definite completion.
//
// Per front j (residual nn, separator na, nJ = na + nn) the reference forms
// M = Sigma(sep, sep), C = chol(M), Y = C^-1 Sigma21, the Schur complement
// S = Sigma11 - Y'Y = R'R (R lower, by reversing the order: the lower
// Cholesky of J S J is J R' J), and returns L21 = -C^-T Y R^-1, D11 = R^-1.
// All of it is one Cholesky of the reordered front matrix
//     K = [ M        Sigma21 J ]
//         [ J Sigma21'  J Sigma11 J ]
// (J reverses the nn residual columns): chol(K) = [C 0; J Y' Ltil] with
// Ltil = chol(J S J), and with X = chol(K)^-1 the outputs are its last nn
// rows: D11(r, c) = X(na + nn-1-c, na + nn-1-r), L21(r, c) = X(na + nn-1-c, r).
// The fronts are independent given Sigma, so the library builds forests of
// these K (plans whose fronts are all roots, in chunks that fit a memory
// budget) and factors them in place (the big ones on the tile-DAG).  A big
// front (nJ > 64) is padded: its separator part grows to na' = 64 ceil(na /
// 64) rows with identity rows and columns between M and the residual part,
// which leaves the Schur complement and hence the outputs unchanged and
// puts the residual part on a tile boundary.  The last nn rows of its
// inverse factor come from backward solves Y = L^-T [0; I_nn] run as tile
// tasks of k_selinv; a small front's factor is inverted here in LDS.
// k_cpl_gather writes K into the forest's panels, k_cpl_extract the
// outputs; both run over one chunk, whose root u is reference front fr[u].
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mf_kernels.h"

namespace mf {

namespace {

// Separator rows of a front's K: na, padded to a multiple of 64 for a big
// front (nn + na > 64).
__device__ __forceinline__ int cpl_sep_rows(int nn, int na) { return nn + na > 64 ? (na + 63) & ~63 : na; }

constexpr int CPL_SEP_LDS = 12288;  // separator rows staged in LDS (ints)

// Workgroups (u, y) for root u (reference front fr[u]): columns c = y,
// y + gridDim.y, ... of its K (nJ' x nJ' column-major at K + koff[u]; the
// lower triangle, which is all the factorization reads).  M(r, c) =
// Sigma(a, b) with a = sep[r] >= b = sep[c]: column b's front o holds it in
// its diagonal block (a in o's residual) or at a's position among o's
// separator rows (sorted), found by binary search in LDS, where o's rows
// are staged once per run of columns owned by o (in global memory when
// there are more than CPL_SEP_LDS).
__global__ __launch_bounds__(256) void k_cpl_gather(DevPlan P, const int32_t* fr, const double* Dval,
                                                    const double* Lval, const int32_t* septgt,
                                                    const int64_t* koff, double* K) {
    __shared__ int32_t srow[CPL_SEP_LDS];
    __shared__ int s_owner;
    const int u = blockIdx.x, tid = threadIdx.x;
    const OutFront o = P.out[fr[u]];
    const int nn = o.nn, na = o.na, np = cpl_sep_rows(nn, na), nj = nn + np;
    double* Kj = K + koff[u];
    const double* D11 = Dval + o.Dp;
    const double* L21 = Lval + o.Lp;
    const int32_t* sep = septgt + o.sepoff;
    if (tid == 0) s_owner = -1;
    __syncthreads();
    for (int c = blockIdx.y; c < nj; c += gridDim.y) {
        double* col = Kj + (int64_t)c * nj;
        if (c < na) {
            // M's column c, rows r >= c (uniform per workgroup)
            const int b = sep[c];
            const int ob = P.col2ref[b];
            const OutFront q = P.out[ob];
            const int lb = b - q.r0;
            const bool staged = q.na <= CPL_SEP_LDS;
            if (staged && ob != s_owner) {
                __syncthreads();  // the previous owner's rows are no longer read
                for (int i = tid; i < q.na; i += blockDim.x) srow[i] = septgt[q.sepoff + i];
                if (tid == 0) s_owner = ob;
                __syncthreads();
            }
            const int32_t* qs = staged ? srow : septgt + q.sepoff;
            for (int r = c + tid; r < na; r += blockDim.x) {
                const int a = sep[r];
                double v;
                if (a < q.r0 + q.nn) {
                    v = Dval[q.Dp + (a - q.r0) + (int64_t)lb * q.nn];
                } else {
                    int lo = 0, hi = q.na;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (qs[mid] < a) lo = mid + 1;
                        else hi = mid;
                    }
                    v = Lval[q.Lp + lo + (int64_t)lb * q.na];
                }
                col[r] = v;
            }
            for (int r = na + tid; r < nj; r += blockDim.x)  // padding rows (0), then J Sigma21'
                col[r] = r < np ? 0.0 : L21[c + (int64_t)(nn - 1 - (r - np)) * na];
        } else if (c < np) {
            for (int r = c + tid; r < nj; r += blockDim.x) col[r] = r == c ? 1.0 : 0.0;  // padding
        } else {
            for (int r = c + tid; r < nj; r += blockDim.x)  // J Sigma11 J
                col[r] = D11[(nn - 1 - (c - np)) + (int64_t)(nn - 1 - (r - np)) * nn];
        }
    }
}

// One workgroup per root u (reference front fr[u]): D11 and L21 of the
// completion from the last nn rows of X = chol(K)^-1.  xoff (NULL: none) with
// xoff[u] >= 0: a big front, Y = (those rows)' (nJ' x nn, ld nJ') at
// X + xoff[u]; else the factor of K (a small front, nJ <= 64) at L + loff[u]
// (ld nJ), inverted here in LDS.
__global__ __launch_bounds__(256) void k_cpl_extract(DevPlan P, const int32_t* fr, const double* X,
                                                     const int64_t* xoff, const double* L,
                                                     const int64_t* loff, double* Dval, double* Lval) {
    __shared__ double Ls[64 * 65], Xs[64 * 65];
    const int u = blockIdx.x, tid = threadIdx.x;
    const OutFront o = P.out[fr[u]];
    const int nn = o.nn, na = o.na, nj = nn + na;
    double* D11 = Dval + o.Dp;
    double* L21 = Lval + o.Lp;
    if (xoff && xoff[u] >= 0) {
        // X(na + q, col) = Y(col', q) with the padded column col' of col
        const int np = cpl_sep_rows(nn, na), ldy = nn + np;
        const double* Y = X + xoff[u];
        for (int e = tid; e < nn * nn; e += blockDim.x) {
            const int c = e / nn, r = e - c * nn;
            D11[e] = r >= c ? Y[(np + nn - 1 - r) + (int64_t)(nn - 1 - c) * ldy] : 0.0;
        }
        for (int e = tid; e < na * nn; e += blockDim.x) {
            const int c = e / na, r = e - c * na;
            L21[e] = Y[r + (int64_t)(nn - 1 - c) * ldy];
        }
        return;
    }
    const double* Xj;
    int64_t ldx;
    {
        // small front: X = L^-1 by forward substitution, column c by thread c
        const double* Lj = L + loff[u];
        for (int e = tid; e < nj * nj; e += blockDim.x) {
            const int c = e / nj, r = e - c * nj;
            Ls[r + c * 65] = r >= c ? Lj[r + (int64_t)c * nj] : 0.0;
        }
        __syncthreads();
        if (tid < nj) {
            const int c = tid;
            for (int r = 0; r < c; ++r) Xs[r + c * 65] = 0.0;
            Xs[c + c * 65] = 1.0 / Ls[c + c * 65];
            for (int r = c + 1; r < nj; ++r) {
                double s = 0.0;
                for (int k = c; k < r; ++k) s += Ls[r + k * 65] * Xs[k + c * 65];
                Xs[r + c * 65] = -s / Ls[r + r * 65];
            }
        }
        __syncthreads();
        Xj = Xs;
        ldx = 65;
    }
    for (int e = tid; e < nn * nn; e += blockDim.x) {
        const int c = e / nn, r = e - c * nn;
        D11[e] = r >= c ? Xj[(na + nn - 1 - c) + (int64_t)(na + nn - 1 - r) * ldx] : 0.0;
    }
    for (int e = tid; e < na * nn; e += blockDim.x) {
        const int c = e / na, r = e - c * na;
        L21[e] = Xj[(na + nn - 1 - c) + (int64_t)r * ldx];
    }
}

}  // namespace

void launch_cpl_gather(const DevPlan& P, int m, const int32_t* fr, const double* Dval, const double* Lval,
                       const int32_t* septgt, const int64_t* koff, double* K, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_cpl_gather, dim3(m, 16), dim3(256), 0, s, P, fr, Dval, Lval, septgt, koff, K);
}

void launch_cpl_extract(const DevPlan& P, int m, const int32_t* fr, const double* X, const int64_t* xoff,
                        const double* L, const int64_t* loff, double* Dval, double* Lval, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_cpl_extract, dim3(m), dim3(256), 0, s, P, fr, X, xoff, L, loff, Dval, Lval);
}

}  // namespace mf
