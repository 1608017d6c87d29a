// Pure-sparse LazyTensor / LazySum application in gather form.
//
// Reference: mul! on a LazySum (src/operators_lazysum.jl:189-238) calls the
// LazyTensor mul! once per term (src/operators_lazytensor.jl:539-609), which
// for all-sparse factors runs the recursive scatter _gemm_recursive_lazy_dense
// (:652-685, left / Ket) or _gemm_recursive_dense_lazy (:613-648, right /
// Bra), each a full read-modify-write pass over the result.
//
// Here every term of the sum is compiled (on the host) into a gather table
// over the term's own sites, and one launch evaluates all terms for each
// output element:
//   out[J] = alpha * ( sum_t sum_{e in row(t, J)} v_e * in[K_e(J)] ) + beta * out[J]
// so the result is written once and the input is read once per term group
// (re-reads are served by LDS/L1/L2, not HBM).  Diagonal terms are folded
// into one coefficient per element before a single multiply with in[J].
//
// k_gather_rows: the gather axis is the fastest (stride 1) axis — a Ket, or
//                the row index of a density matrix (left application).
// k_gather_cols: the gather axis is the column index of a density matrix
//                (right application); lanes run along the contiguous rows and
//                the whole term walk is wave-uniform.
#include "qob_internal.h"

#include <algorithm>

namespace qob {

namespace {

template <typename T> __device__ __forceinline__ T to_t(double2 a) {
    T r; r.x = (typename real_of<T>::type)a.x; r.y = (typename real_of<T>::type)a.y; return r;
}

struct Digits {
    // non-square only: r-space offset of the element and mask of sites whose
    // l-digit is outside the r-range (identity there contributes nothing)
    int64_t jr_all;
    uint32_t invalid;
};

__device__ __forceinline__ uint32_t digit_of(uint32_t J, const int32_t *s) {
    // s: stride_l, m_stride, l_stride, d_l, m_d, l_d
    const uint32_t qv = fdiv(J, (uint32_t)s[1], (uint32_t)s[2]);
    return qv - (uint32_t)s[3] * fdiv(qv, (uint32_t)s[4], (uint32_t)s[5]);
}

__device__ __forceinline__ Digits all_digits(uint32_t J, int nsites, const int32_t *__restrict__ site_info) {
    Digits d;
    d.jr_all = 0;
    d.invalid = 0;
    for (int s = 0; s < nsites; ++s) {
        const int32_t *si = site_info + 8 * s;
        const uint32_t j = digit_of(J, si);
        d.jr_all += (int64_t)j * si[6];
        if (j >= (uint32_t)si[7]) d.invalid |= (1u << s);
    }
    return d;
}

// Sum over all terms for output J.  Returns (acc, dacc): gathered part and
// diagonal coefficient.  LD(K) loads in[K].
template <typename T, bool SQUARE, typename LD>
__device__ __forceinline__ T eval_terms(uint32_t J, const GatherPlanDev &P, const Digits &dg, LD ld) {
    const T *__restrict__ vals = (const T *)P.vals;
    T acc = czero<T>();
    T dacc = czero<T>();
    const int32_t *pg = P.prog;
    for (int t = 0; t < P.nterms; ++t) {
        const int kind = pg[0], flags = pg[1], q = pg[2], base = pg[3];
        const uint32_t fmask = (uint32_t)pg[4];
        const int32_t *sr = pg + GP_HDR;
        pg += GP_HDR + GP_SITE * q;
        if (!SQUARE && (dg.invalid & ~fmask)) continue;
        int64_t kb = SQUARE ? (int64_t)J : dg.jr_all;
        if (kind == 0) {
            int jloc = 0;
            for (int i = 0; i < q; ++i) {
                const int32_t *s = sr + GP_SITE * i;
                const uint32_t j = digit_of(J, s);
                jloc += (int)j * s[6];
                if (!SQUARE) kb -= (int64_t)j * s[8];
            }
            const int32_t p0 = P.rowptr[base + jloc], p1 = P.rowptr[base + jloc + 1];
            if (flags & 1) {
                for (int32_t p = p0; p < p1; ++p) dacc = cadd(dacc, vals[p]);
            } else {
                for (int32_t p = p0; p < p1; ++p) acc = cfma(vals[p], ld(kb + P.koff[p]), acc);
            }
        } else {
            T v = mk((typename real_of<T>::type)1, (typename real_of<T>::type)0);
            int64_t off = 0;
            bool live = true;
            for (int i = 0; i < q; ++i) {
                const int32_t *s = sr + GP_SITE * i;
                const uint32_t j = digit_of(J, s);
                const int e = s[7] + (int)j;
                const T a = vals[e];
                live = live && (a.x != 0 || a.y != 0);
                v = cmul(v, a);
                off += P.koff[e];
                if (!SQUARE) kb -= (int64_t)j * s[8];
            }
            if (live) {
                if (flags & 1) dacc = cadd(dacc, v);
                else acc = cfma(v, ld(kb + off), acc);
            }
        }
    }
    if (dacc.x != 0 || dacc.y != 0) {
        // diagonal terms only exist for square operators
        acc = cfma(dacc, ld((int64_t)J), acc);
    }
    return acc;
}

template <typename T, bool SQUARE>
__global__ void __launch_bounds__(256) k_gather_rows(GatherPlanDev P, const T *__restrict__ in, int64_t n_in,
                                                     T *__restrict__ out, int64_t n_out, int64_t nbatch,
                                                     double2 alpha_, double2 beta_, int beta_zero) {
    const int64_t J = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (J >= n_out) return;
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    Digits dg{0, 0};
    if (!SQUARE) dg = all_digits((uint32_t)J, P.nsites, P.site_info);
    for (int64_t b = blockIdx.y; b < nbatch; b += gridDim.y) {
        const T *ib = in + b * n_in;
        T *o = out + b * n_out + J;
        const T acc = eval_terms<T, SQUARE>((uint32_t)J, P, dg, [&](int64_t K) { return ib[K]; });
        *o = axpby(alpha, acc, beta, o, beta_zero != 0);
    }
}

template <typename T, bool SQUARE>
__global__ void __launch_bounds__(256) k_gather_cols(GatherPlanDev P, const T *__restrict__ in, int64_t n_in,
                                                     T *__restrict__ out, int64_t n_out, int64_t nrows,
                                                     double2 alpha_, double2 beta_, int beta_zero) {
    (void)n_in;  // the plan is compiled for this input length (the entry points check it)
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    const bool ok = i < nrows;
    for (int64_t J = blockIdx.y; J < n_out; J += gridDim.y) {
        Digits dg{0, 0};
        if (!SQUARE) dg = all_digits((uint32_t)J, P.nsites, P.site_info);
        const T *ib = in + (ok ? i : 0);
        const T acc = eval_terms<T, SQUARE>((uint32_t)J, P, dg, [&](int64_t K) { return ib[K * nrows]; });
        if (ok) {
            T *o = out + i + J * nrows;
            *o = axpby(alpha, acc, beta, o, beta_zero != 0);
        }
    }
}

// FP64 block sum of a double2 (256 threads); lane 0 of wave 0 gets the total
__device__ __forceinline__ double2 block_sum256(double2 v) {
    __shared__ double2 sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[threadIdx.x].x += sh[threadIdx.x + w].x;
            sh[threadIdx.x].y += sh[threadIdx.x + w].y;
        }
        __syncthreads();
    }
    return sh[0];
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_gather_dot(GatherPlanDev P, const T *__restrict__ u, const T *__restrict__ v,
                                                    int64_t n, double2 *__restrict__ partials) {
    double2 r = make_double2(0, 0);
    for (int64_t J = blockIdx.x * (int64_t)256 + threadIdx.x; J < n; J += (int64_t)gridDim.x * 256) {
        Digits dg{0, 0};
        T acc;
        if (MODE == 0) {
            acc = eval_terms<T, true>((uint32_t)J, P, dg, [&](int64_t K) { return v[K]; });
            const T a = u[J];
            r.x += (double)a.x * acc.x + (double)a.y * acc.y;
            r.y += (double)a.x * acc.y - (double)a.y * acc.x;
        } else {
            const T *col = v + J * n;
            acc = eval_terms<T, true>((uint32_t)J, P, dg, [&](int64_t K) { return col[K]; });
            r.x += acc.x;
            r.y += acc.y;
        }
    }
    const double2 t = block_sum256(r);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

}  // namespace

int gather_dot_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256)); }

hipError_t launch_gather_dot(qob_dtype dt, const GatherPlanDev &P, int mode, const void *u, const void *v, int64_t n,
                             double2 *partials, hipStream_t s) {
    if (!P.square) return hipErrorInvalidValue;
    const int g = gather_dot_blocks(n);
#define QOB_GD(TT, MD) \
    hipLaunchKernelGGL((k_gather_dot<TT, MD>), dim3(g), dim3(256), 0, s, P, (const TT *)u, (const TT *)v, n, partials)
    if (dt == QOB_C64) { if (mode == 0) QOB_GD(double2, 0); else QOB_GD(double2, 1); }
    else { if (mode == 0) QOB_GD(float2, 0); else QOB_GD(float2, 1); }
#undef QOB_GD
    return hipGetLastError();
}

hipError_t launch_gather_rows(qob_dtype dt, const GatherPlanDev &P, const void *in, int64_t n_in, void *out,
                              int64_t n_out, int64_t nbatch, double2 alpha, double2 beta, hipStream_t s) {
    if (n_out == 0 || nbatch == 0) return hipSuccess;
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    const dim3 g((uint32_t)((n_out + 255) / 256), (uint32_t)(nbatch < 65535 ? nbatch : 65535));
#define QOB_GR(TT, SQ) \
    hipLaunchKernelGGL((k_gather_rows<TT, SQ>), g, dim3(256), 0, s, P, (const TT *)in, n_in, (TT *)out, n_out, nbatch, alpha, beta, bz)
    if (dt == QOB_C64) { if (P.square) QOB_GR(double2, true); else QOB_GR(double2, false); }
    else { if (P.square) QOB_GR(float2, true); else QOB_GR(float2, false); }
#undef QOB_GR
    return hipGetLastError();
}

hipError_t launch_gather_cols(qob_dtype dt, const GatherPlanDev &P, const void *in, int64_t n_in, void *out,
                              int64_t n_out, int64_t nrows, double2 alpha, double2 beta, hipStream_t s) {
    if (n_out == 0 || nrows == 0) return hipSuccess;
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    (void)n_in;
    const dim3 g((uint32_t)((nrows + 255) / 256), (uint32_t)(n_out < 65535 ? n_out : 65535));
#define QOB_GC(TT, SQ) \
    hipLaunchKernelGGL((k_gather_cols<TT, SQ>), g, dim3(256), 0, s, P, (const TT *)in, n_in, (TT *)out, n_out, nrows, alpha, beta, bz)
    if (dt == QOB_C64) { if (P.square) QOB_GC(double2, true); else QOB_GC(double2, false); }
    else { if (P.square) QOB_GC(float2, true); else QOB_GC(float2, false); }
#undef QOB_GC
    return hipGetLastError();
}

}  // namespace qob
