// Elementwise kernels and FP64 reductions.
//   scale  : _zero_op_mul! / rmul! (src/operators_lazysum.jl:179-186,
//            src/sparsematrix.jl:103-107)
//   axpy   : the Ket update of a time step (y += a x)
//   dot    : dot(x, y) = sum conj(x) y (src/states.jl:80 with dagger)
//   nrm2   : norm (src/states.jl:89)
//   trace  : tr (src/operators_dense.jl:189)
//   refresh: plan values as linear combinations of device-held LazySum
//            coefficients (update_static_coefficients!,
//            src/time_dependent_operator.jl:286-297, done on the device)
// Reductions accumulate in double for both element types and run in two
// deterministic stages (per-block partials, then one block), so a result
// does not depend on scheduling.
#include "qob_internal.h"

#include <atomic>

namespace qob {

namespace {

constexpr int BS = 256;

template <typename T> __device__ __forceinline__ T from_d2(double2 a) {
    T r; r.x = (typename real_of<T>::type)a.x; r.y = (typename real_of<T>::type)a.y; return r;
}

template <typename T>
__global__ void __launch_bounds__(BS) k_scale(int64_t n, double2 a_, T *x) {
    const T a = from_d2<T>(a_);
    const bool zero = (a_.x == 0.0 && a_.y == 0.0);
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        x[i] = zero ? czero<T>() : cmul(a, x[i]);
    }
}

// slot s = sum_{p in [ptr[s], ptr[s+1])} bval[p] * coef[term[p]], in FP64;
// slots below nvals are plan values of element type T, the others complex
// doubles written into the program at word offsets dst[s - nvals]
template <typename T>
__global__ void __launch_bounds__(BS) k_refresh(int nslots, int nvals, const int32_t *__restrict__ ptr,
                                                const int32_t *__restrict__ term, const double2 *__restrict__ bval,
                                                const int32_t *__restrict__ dst, const double2 *__restrict__ coef,
                                                T *__restrict__ vals, int32_t *__restrict__ prog) {
    const int s = blockIdx.x * BS + threadIdx.x;
    if (s >= nslots) return;
    double re = 0, im = 0;
    for (int p = ptr[s]; p < ptr[s + 1]; ++p) {
        const double2 b = bval[p], c = coef[term[p]];
        re += b.x * c.x - b.y * c.y;
        im += b.x * c.y + b.y * c.x;
    }
    if (s < nvals) {
        vals[s] = from_d2<T>(make_double2(re, im));
    } else {
        *(double2 *)(prog + dst[s - nvals]) = make_double2(re, im);
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_copy(int64_t n, const T *__restrict__ x, T *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) y[i] = x[i];
}

// 4 elements per lane in flight, non-temporal stores
template <typename T>
__global__ void __launch_bounds__(BS) k_copy4(int64_t n, const T *__restrict__ x, T *__restrict__ y) {
    const int64_t base = blockIdx.x * (int64_t)BS * 4 + threadIdx.x;
    T v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + (int64_t)k * BS;
        if (i < n) v[k] = x[i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + (int64_t)k * BS;
        if (i < n) st_stream(y + i, v[k]);
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_axpy(int64_t n, double2 a_, const T *__restrict__ x, T *__restrict__ y) {
    const T a = from_d2<T>(a_);
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        y[i] = cfma(a, x[i], y[i]);
    }
}

// y = a x + b y  (b == 0 discards y)
template <typename T>
__global__ void __launch_bounds__(BS) k_axpby(int64_t n, double2 a_, const T *__restrict__ x, double2 b_, T *__restrict__ y) {
    const T a = from_d2<T>(a_), b = from_d2<T>(b_);
    const bool bz = (b_.x == 0.0 && b_.y == 0.0);
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        T r = cmul(a, x[i]);
        if (!bz) r = cfma(b, y[i], r);
        y[i] = r;
    }
}

template <typename S, typename D>
__global__ void __launch_bounds__(BS) k_convert(int64_t n, const S *__restrict__ x, D *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        S v = x[i];
        D r; r.x = (typename real_of<D>::type)v.x; r.y = (typename real_of<D>::type)v.y;
        y[i] = r;
    }
}

__device__ __forceinline__ double2 block_sum(double2 v, double2 *sh) {
    // wave reduction then LDS across the 4 waves of the block
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_down(v.x, off, 64);
        v.y += __shfl_down(v.y, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double2 r = make_double2(0, 0);
    if (threadIdx.x == 0) {
        for (int i = 0; i < BS / 64; ++i) { r.x += sh[i].x; r.y += sh[i].y; }
    }
    return r;
}

// mode 0: dot (conj(x) y), 1: |x|^2 in .x, 2: trace (x = A, stride ld+1)
template <typename T, int MODE>
__global__ void __launch_bounds__(BS) k_reduce1(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                               int64_t step, double2 *__restrict__ partials) {
    __shared__ double2 sh[BS / 64];
    double2 acc = make_double2(0, 0);
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        if (MODE == 0) {
            T a = x[i], b = y[i];
            double ax = a.x, ay = a.y, bx = b.x, by = b.y;
            acc.x = fma(ax, bx, fma(ay, by, acc.x));
            acc.y = fma(ax, by, fma(-ay, bx, acc.y));
        } else if (MODE == 1) {
            T a = x[i];
            double ax = a.x, ay = a.y;
            acc.x = fma(ax, ax, fma(ay, ay, acc.x));
        } else {
            T a = x[i * step];
            acc.x += (double)a.x;
            acc.y += (double)a.y;
        }
    }
    double2 r = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// final stage: one block sums nparts partials; mode 1 writes sqrt into a double
template <int MODE>
__global__ void __launch_bounds__(BS) k_reduce2(int nparts, const double2 *__restrict__ partials, void *out) {
    __shared__ double2 sh[BS / 64];
    double2 acc = make_double2(0, 0);
    for (int i = threadIdx.x; i < nparts; i += BS) {
        acc.x += partials[i].x;
        acc.y += partials[i].y;
    }
    double2 r = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        if (MODE == 1) *(double *)out = sqrt(r.x);
        else *(double2 *)out = r;
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_scale_inv_norm(int64_t n, const double *__restrict__ nrm, T *x) {
    const double inv = 1.0 / *nrm;
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        T v = x[i];
        v.x = (typename real_of<T>::type)(v.x * inv);
        v.y = (typename real_of<T>::type)(v.y * inv);
        x[i] = v;
    }
}

// A /= tr  (complex division, FP64)
template <typename T>
__global__ void __launch_bounds__(BS) k_scale_inv_trace(int64_t n, const double2 *__restrict__ tr, T *x) {
    const double2 t = *tr;
    const double den = t.x * t.x + t.y * t.y;
    const double ix = t.x / den, iy = -t.y / den;
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        T v = x[i];
        double vx = v.x, vy = v.y;
        v.x = (typename real_of<T>::type)(vx * ix - vy * iy);
        v.y = (typename real_of<T>::type)(vx * iy + vy * ix);
        x[i] = v;
    }
}

// out (ro x co) = alpha * E(in) + beta * out where E zero-pads / truncates
// in (ri x ci) to ro x co: the action of Eye(m, n) from the left (rows) or the
// right (columns) (FillArrays Eye as used by identityoperator,
// src/operators_sparse.jl:180-183).
template <typename T>
__global__ void __launch_bounds__(BS) k_embed(int64_t ro, int64_t co, const T *__restrict__ in, int64_t ri, int64_t ci,
                                              T *__restrict__ out, double2 a_, double2 b_, int bz) {
    const T a = from_d2<T>(a_), b = from_d2<T>(b_);
    const int64_t total = ro * co;
    for (int64_t k = blockIdx.x * (int64_t)BS + threadIdx.x; k < total; k += (int64_t)gridDim.x * BS) {
        const int64_t j = k / ro, i = k - j * ro;
        T v = (i < ri && j < ci) ? in[i + j * ri] : czero<T>();
        T r = cmul(a, v);
        if (!bz) r = cfma(b, out[k], r);
        out[k] = r;
    }
}

inline unsigned grid_for(int64_t n) {
    int64_t g = (n + BS - 1) / BS;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

hipError_t launch_embed(qob_dtype dt, const void *in, int64_t ri, int64_t ci, void *out, int64_t ro, int64_t co,
                        double2 a, double2 b, hipStream_t s) {
    const int64_t total = ro * co;
    if (total <= 0) return hipSuccess;
    const int bz = (b.x == 0.0 && b.y == 0.0);
    if (dt == QOB_C64)
        hipLaunchKernelGGL(k_embed<double2>, grid_for(total), BS, 0, s, ro, co, (const double2 *)in, ri, ci, (double2 *)out, a, b, bz);
    else
        hipLaunchKernelGGL(k_embed<float2>, grid_for(total), BS, 0, s, ro, co, (const float2 *)in, ri, ci, (float2 *)out, a, b, bz);
    return hipGetLastError();
}

int g_copy_variant = 1;

int reduce_blocks(int64_t n) {
    int64_t g = (n + BS - 1) / BS;
    if (g > 1024) g = 1024;
    if (g < 1) g = 1;
    return (int)g;
}

int device_cu_count() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cache[dev].load(std::memory_order_relaxed);
    if (n > 0) return n;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cache[dev].store(n, std::memory_order_relaxed);
    return n;
}

hipError_t launch_scale(qob_dtype dt, int64_t n, double2 a, void *x, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (dt == QOB_C64) hipLaunchKernelGGL(k_scale<double2>, grid_for(n), BS, 0, s, n, a, (double2 *)x);
    else hipLaunchKernelGGL(k_scale<float2>, grid_for(n), BS, 0, s, n, a, (float2 *)x);
    return hipGetLastError();
}

hipError_t launch_refresh(qob_dtype dt, int nslots, int nvals, const int32_t *ptr, const int32_t *term,
                          const double2 *bval, const int32_t *dst, const double2 *coef, void *vals, int32_t *prog,
                          hipStream_t s) {
    if (nslots <= 0) return hipSuccess;
    const unsigned g = (unsigned)((nslots + BS - 1) / BS);
    if (dt == QOB_C64)
        hipLaunchKernelGGL(k_refresh<double2>, dim3(g), dim3(BS), 0, s, nslots, nvals, ptr, term, bval, dst, coef,
                           (double2 *)vals, prog);
    else
        hipLaunchKernelGGL(k_refresh<float2>, dim3(g), dim3(BS), 0, s, nslots, nvals, ptr, term, bval, dst, coef,
                           (float2 *)vals, prog);
    return hipGetLastError();
}

hipError_t launch_copy(qob_dtype dt, int64_t n, const void *x, void *y, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (g_copy_variant == 1 && (n + 4 * BS - 1) / (4 * BS) < 0x7fffffff) {
        const unsigned g4 = (unsigned)((n + 4 * BS - 1) / (4 * BS));
        if (dt == QOB_C64) hipLaunchKernelGGL(k_copy4<double2>, g4, BS, 0, s, n, (const double2 *)x, (double2 *)y);
        else hipLaunchKernelGGL(k_copy4<float2>, g4, BS, 0, s, n, (const float2 *)x, (float2 *)y);
        return hipGetLastError();
    }
    const unsigned g = (unsigned)std::min<int64_t>((n + BS - 1) / BS, 1 << 20);
    if (dt == QOB_C64) hipLaunchKernelGGL(k_copy<double2>, g, BS, 0, s, n, (const double2 *)x, (double2 *)y);
    else hipLaunchKernelGGL(k_copy<float2>, g, BS, 0, s, n, (const float2 *)x, (float2 *)y);
    return hipGetLastError();
}

hipError_t launch_axpy(qob_dtype dt, int64_t n, double2 a, const void *x, void *y, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (dt == QOB_C64) hipLaunchKernelGGL(k_axpy<double2>, grid_for(n), BS, 0, s, n, a, (const double2 *)x, (double2 *)y);
    else hipLaunchKernelGGL(k_axpy<float2>, grid_for(n), BS, 0, s, n, a, (const float2 *)x, (float2 *)y);
    return hipGetLastError();
}

hipError_t launch_axpby_copy(qob_dtype dt, int64_t n, double2 a, const void *x, double2 b, void *y, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (dt == QOB_C64) hipLaunchKernelGGL(k_axpby<double2>, grid_for(n), BS, 0, s, n, a, (const double2 *)x, b, (double2 *)y);
    else hipLaunchKernelGGL(k_axpby<float2>, grid_for(n), BS, 0, s, n, a, (const float2 *)x, b, (float2 *)y);
    return hipGetLastError();
}

hipError_t launch_convert(qob_dtype from, const void *src, qob_dtype to, void *dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (from == QOB_C64 && to == QOB_C32)
        hipLaunchKernelGGL((k_convert<double2, float2>), grid_for(n), BS, 0, s, n, (const double2 *)src, (float2 *)dst);
    else if (from == QOB_C32 && to == QOB_C64)
        hipLaunchKernelGGL((k_convert<float2, double2>), grid_for(n), BS, 0, s, n, (const float2 *)src, (double2 *)dst);
    else if (from == QOB_C64)
        return hipMemcpyAsync(dst, src, (size_t)n * 16, hipMemcpyDeviceToDevice, s);
    else
        return hipMemcpyAsync(dst, src, (size_t)n * 8, hipMemcpyDeviceToDevice, s);
    return hipGetLastError();
}

template <int MODE>
static hipError_t reduce_any(qob_dtype dt, int64_t n, const void *x, const void *y, int64_t step,
                             double2 *partials, void *out, hipStream_t s) {
    const int g = reduce_blocks(n);
    if (dt == QOB_C64)
        hipLaunchKernelGGL((k_reduce1<double2, MODE>), g, BS, 0, s, n, (const double2 *)x, (const double2 *)y, step, partials);
    else
        hipLaunchKernelGGL((k_reduce1<float2, MODE>), g, BS, 0, s, n, (const float2 *)x, (const float2 *)y, step, partials);
    hipLaunchKernelGGL(k_reduce2<MODE>, 1, BS, 0, s, g, (const double2 *)partials, out);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(int nparts, const double2 *partials, double2 *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce2<0>, 1, BS, 0, s, nparts, partials, (void *)out);
    return hipGetLastError();
}

hipError_t launch_dot(qob_dtype dt, int64_t n, const void *x, const void *y, double2 *partials, double2 *out, hipStream_t s) {
    return reduce_any<0>(dt, n, x, y, 1, partials, out, s);
}
hipError_t launch_nrm2(qob_dtype dt, int64_t n, const void *x, double2 *partials, double *out, hipStream_t s) {
    return reduce_any<1>(dt, n, x, x, 1, partials, out, s);
}
hipError_t launch_trace(qob_dtype dt, int64_t n, const void *A, double2 *partials, double2 *out, hipStream_t s) {
    return reduce_any<2>(dt, n, A, A, n + 1, partials, out, s);
}

hipError_t launch_scale_by_inv_norm(qob_dtype dt, int64_t n, const double *nrm, void *x, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (dt == QOB_C64) hipLaunchKernelGGL(k_scale_inv_norm<double2>, grid_for(n), BS, 0, s, n, nrm, (double2 *)x);
    else hipLaunchKernelGGL(k_scale_inv_norm<float2>, grid_for(n), BS, 0, s, n, nrm, (float2 *)x);
    return hipGetLastError();
}

hipError_t launch_scale_by_inv_trace(qob_dtype dt, int64_t n, int64_t ld, const double2 *tr, void *A, hipStream_t s) {
    const int64_t total = n * ld;
    if (total <= 0) return hipSuccess;
    if (dt == QOB_C64) hipLaunchKernelGGL(k_scale_inv_trace<double2>, grid_for(total), BS, 0, s, total, tr, (double2 *)A);
    else hipLaunchKernelGGL(k_scale_inv_trace<float2>, grid_for(total), BS, 0, s, total, tr, (float2 *)A);
    return hipGetLastError();
}

}  // namespace qob
