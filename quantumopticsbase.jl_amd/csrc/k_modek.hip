// Dense mode-k product: the unit step of the LazyTensor tensor-contraction
// engine.
//
// Reference: _tp_matmul! (src/operators_lazytensor.jl:406-428) dispatches to
// _tp_matmul_first! (:281-290, a GEMM on a reshape), _tp_matmul_last!
// (:292-301) or _tp_matmul_mid! (:333-404), which permutes the middle mode to
// the front or back through cached temporaries, runs a GEMM and permutes
// back.  On the GPU no permutation is materialised: the kernel reads the
// strided fibres of the input directly,
//   out[i1, a, i3] = alpha * sum_b A[a, b] * in[i1, b, i3] + beta * out[i1, a, i3]
// with in viewed as (n1, d_in, n3) and out as (n1, d_out, n3) in column-major
// order.  One lane owns one fibre (i1, i3): it loads the d_in inputs once into
// registers and writes d_out outputs.  Consecutive lanes take consecutive i1,
// so for every mode but the first the loads and stores are coalesced; for
// the first mode (n1 == 1) each lane reads one contiguous fibre.  A is read
// with wave-uniform indices (scalar loads / broadcast).
#include "qob_internal.h"

namespace qob {

namespace {

template <typename T> __device__ __forceinline__ T to_t(double2 a) {
    T r; r.x = (typename real_of<T>::type)a.x; r.y = (typename real_of<T>::type)a.y; return r;
}

template <typename T, int MAXD>
__global__ void __launch_bounds__(512) k_modek(const T *__restrict__ A, int d_out, int d_in,
                                               const T *__restrict__ in, T *__restrict__ out, int64_t n1,
                                               int64_t n3, double2 alpha_, double2 beta_, int beta_zero) {
    const int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (f >= n1 * n3) return;
    const int64_t i3 = f / n1, i1 = f - i3 * n1;
    const T *ib = in + i1 + i3 * n1 * d_in;
    T *ob = out + i1 + i3 * n1 * d_out;
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    T x[MAXD];
#pragma unroll
    for (int b = 0; b < MAXD; ++b) x[b] = (b < d_in) ? ib[(int64_t)b * n1] : czero<T>();
    for (int a = 0; a < d_out; ++a) {
        T acc = czero<T>();
#pragma unroll
        for (int b = 0; b < MAXD; ++b) {
            if (b < d_in) acc = cfma(A[a + (int64_t)b * d_out], x[b], acc);
        }
        T *o = ob + (int64_t)a * n1;
        *o = axpby(alpha, acc, beta, o, beta_zero != 0);
    }
}

}  // namespace

int g_modek_nth = 64;  // fibres (lanes) per workgroup: 64, 128, 256 or 512 (64 measured fastest)

bool modek_supported(int64_t d_out, int64_t d_in) { return d_in <= 64 && d_out >= 1 && d_in >= 1; }

hipError_t launch_modek(qob_dtype dt, const void *A, int64_t d_out, int64_t d_in, const void *in, void *out,
                        int64_t n1, int64_t n3, double2 alpha, double2 beta, hipStream_t s) {
    const int64_t nf = n1 * n3;
    if (nf == 0 || d_out == 0) return hipSuccess;
    if (!modek_supported(d_out, d_in)) return hipErrorInvalidValue;
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    const int nth = g_modek_nth;
    const uint32_t g = (uint32_t)((nf + nth - 1) / nth);
#define QOB_MK(TT, MD)                                                                                     \
    hipLaunchKernelGGL((k_modek<TT, MD>), dim3(g), dim3(nth), 0, s, (const TT *)A, (int)d_out, (int)d_in, \
                       (const TT *)in, (TT *)out, n1, n3, alpha, beta, bz)
    if (dt == QOB_C64) {
        if (d_in <= 4) QOB_MK(double2, 4);
        else if (d_in <= 16) QOB_MK(double2, 16);
        else if (d_in <= 32) QOB_MK(double2, 32);
        else QOB_MK(double2, 64);
    } else {
        if (d_in <= 4) QOB_MK(float2, 4);
        else if (d_in <= 16) QOB_MK(float2, 16);
        else if (d_in <= 32) QOB_MK(float2, 32);
        else QOB_MK(float2, 64);
    }
#undef QOB_MK
    return hipGetLastError();
}

}  // namespace qob
