// MFMA helpers shared by the dense kernels (k_fused.hip, k_mode.hip).
//
// v_mfma_{f64,f32}_16x16x4: D (16 x 16) += A (16 x 4) * B (4 x 16), one wave.
// Lane layouts (checked on gfx950 with a one-wave probe):
//   A-frag  A[i][k]  lane i + 16 k          (i < 16, k < 4)
//   B-frag  B[k][j]  lane j + 16 k
//   f32 D   D[4 (lane/16) + r][lane % 16]   register r
//   f64 D   D[(lane/16) + 4 r][lane % 16]   register r
#pragma once

#include "qob_internal.h"

namespace qob {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename R> struct mf;
template <> struct mf<double> {
    using v4 = v4d;
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of the D tile held by register r of this lane
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct mf<float> {
    using v4 = v4f;
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// One complex k-step of a tile: (re, im) += a * b with a an A-fragment value
// and b a B-fragment value, formed from four real products as BLAS forms a
// complex product (re += ar br - ai bi, im += ar bi + ai br).
template <typename R>
__device__ __forceinline__ void cmma4(cplx<R> a, cplx<R> b, typename mf<R>::v4 &re, typename mf<R>::v4 &im) {
    re = mf<R>::mma(a.x, b.x, re);
    re = mf<R>::mma(-a.y, b.y, re);
    im = mf<R>::mma(a.x, b.y, im);
    im = mf<R>::mma(a.y, b.x, im);
}

}  // namespace qob
