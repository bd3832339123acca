// mf_mfma.h — the gfx950 matrix-core operation used by the dense tile
// products of the factorization and of the solves (device code only).
#pragma once

#include <hip/hip_runtime.h>

namespace mf {

// v_mfma_*_16x16x4: lane l holds A[l%16][l/16] and B[l/16][l%16]; output
// register v of lane l holds column l%16 and row Mfma<T>::row(l/16, v).
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

}  // namespace mf
