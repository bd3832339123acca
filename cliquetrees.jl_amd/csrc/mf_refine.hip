// mf_refine.hip — FP64 kernels of iterative refinement around a factor of
// any front precision: residual R = B - A X with the symmetric matrix in
// full CSC form (1-based, original ordering), squared norms per column and
// row segment (one segment per system of a batch) and the update X += D.
// All are HBM-bound.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mf_kernels.h"

namespace mf {

namespace {

// One thread per (row, right-hand side); rows are consecutive within a
// wavefront.  Column i of the symmetric A is row i.
__global__ __launch_bounds__(256) void k_residual(int64_t n, const int64_t* colptr,
                                                  const int64_t* rowval, const double* val,
                                                  const double* X, int64_t ldx, const double* B,
                                                  int64_t ldb, double* R, int64_t ldr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= n) return;
    const double* x = X + j * ldx;
    double s = 0.0;
    const int64_t p0 = colptr[i] - 1, p1 = colptr[i + 1] - 1;
    // eight nonzeros at a time, every load issued before the sums (indices
    // past the row's end read its first entry); the sum runs in p order
    for (int64_t pb = p0; pb < p1; pb += 8) {
        int64_t r[8];
        double v[8], xv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t pk = pb + k < p1 ? pb + k : p0;
            r[k] = rowval[pk];
            v[k] = val[pk];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) xv[k] = x[r[k] - 1];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (pb + k < p1) s += v[k] * xv[k];
    }
    R[i + j * ldr] = B[i + j * ldb] - s;
}

// out[j * nseg + g] = sum over rows [seg[g], seg[g+1]) of M(i, j)^2; one
// workgroup per (segment, column), fixed reduction order.
__global__ __launch_bounds__(256) void k_colnorm2(const int64_t* seg, int nseg, const double* M,
                                                  int64_t ld, double* out) {
    __shared__ double part[256];
    const int g = blockIdx.x, j = blockIdx.y, tid = threadIdx.x;
    const double* m = M + j * ld;
    double s = 0.0;
    for (int64_t i = seg[g] + tid; i < seg[g + 1]; i += 256) s += m[i] * m[i];
    part[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) part[tid] += part[tid + w];
        __syncthreads();
    }
    if (tid == 0) out[(int64_t)j * nseg + g] = part[0];
}

__global__ __launch_bounds__(256) void k_axpy(int64_t n, const double* D, int64_t ldd, double* X,
                                              int64_t ldx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) X[i + j * ldx] += D[i + j * ldd];
}

__global__ __launch_bounds__(256) void k_copy_cols(int64_t n, const double* S, int64_t lds,
                                                   double* D, int64_t ldd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) D[i + j * ldd] = S[i + j * lds];
}

}  // namespace

void launch_residual(int64_t n, const int64_t* colptr, const int64_t* rowval, const double* val,
                     const double* X, int64_t ldx, const double* B, int64_t ldb, double* R,
                     int64_t ldr, int nrhs, hipStream_t s) {
    if (n <= 0 || nrhs <= 0) return;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)nrhs);
    hipLaunchKernelGGL(k_residual, grid, dim3(256), 0, s, n, colptr, rowval, val, X, ldx, B, ldb,
                       R, ldr);
}

void launch_colnorm2(const int64_t* seg, int nseg, const double* M, int64_t ld, int nrhs,
                     double* out, hipStream_t s) {
    if (nrhs <= 0 || nseg <= 0) return;
    hipLaunchKernelGGL(k_colnorm2, dim3(nseg, nrhs), dim3(256), 0, s, seg, nseg, M, ld, out);
}

void launch_axpy(int64_t n, const double* D, int64_t ldd, double* X, int64_t ldx, int nrhs,
                 hipStream_t s) {
    if (n <= 0 || nrhs <= 0) return;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)nrhs);
    hipLaunchKernelGGL(k_axpy, grid, dim3(256), 0, s, n, D, ldd, X, ldx);
}

void launch_copy_cols(int64_t n, const double* S, int64_t lds, double* D, int64_t ldd, int nrhs,
                      hipStream_t s) {
    if (n <= 0 || nrhs <= 0) return;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)nrhs);
    hipLaunchKernelGGL(k_copy_cols, grid, dim3(256), 0, s, n, S, lds, D, ldd);
}

}  // namespace mf
