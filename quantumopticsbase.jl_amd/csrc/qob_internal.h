// Internal declarations shared by the HIP kernels and the host API of
// libqob_hip.so.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <algorithm>

#include "../../include/qob_hip.h"

namespace qob {

// ------------------------------------------------------------ complex math
// Interleaved (re, im) pairs, the layout of Julia's ComplexF64/ComplexF32.
template <typename R> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename R> using cplx = typename cplx_t<R>::type;

template <typename T> struct real_of;
template <> struct real_of<double2> { using type = double; };
template <> struct real_of<float2> { using type = float; };

__host__ __device__ inline double2 mk(double re, double im) { return make_double2(re, im); }
__host__ __device__ inline float2 mk(float re, float im) { return make_float2(re, im); }

template <typename T> __host__ __device__ inline T czero() { T z; z.x = 0; z.y = 0; return z; }

// a * b
template <typename T> __device__ __forceinline__ T cmul(T a, T b) {
    T r;
    r.x = fma(a.x, b.x, -a.y * b.y);
    r.y = fma(a.x, b.y, a.y * b.x);
    return r;
}
// c + a * b
template <typename T> __device__ __forceinline__ T cfma(T a, T b, T c) {
    T r;
    r.x = fma(a.x, b.x, fma(-a.y, b.y, c.x));
    r.y = fma(a.x, b.y, fma(a.y, b.x, c.y));
    return r;
}
template <typename T> __device__ __forceinline__ T cadd(T a, T b) { T r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T> __device__ __forceinline__ T cconj(T a) { a.y = -a.y; return a; }

// alpha * acc + beta * old, with beta == 0 discarding old (NaN-safe).
template <typename T> __device__ __forceinline__ T axpby(T alpha, T acc, T beta, const T *old, bool beta_zero) {
    T r = cmul(alpha, acc);
    if (!beta_zero) r = cfma(beta, *old, r);
    return r;
}

// Streaming store of a result element that is not read again by the same
// kernel: the non-temporal hint keeps it from evicting input lines that
// later gathers of the same launch still need.
template <typename T> __device__ __forceinline__ void st_stream(T *p, T v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}
template <typename T> __device__ __forceinline__ void st_out(T *p, T v, bool nt) {
    if (nt) st_stream(p, v);
    else *p = v;
}

// Accesses through pointers that a kernel reads from a device table (or a
// struct argument): the compiler cannot tell that they point to global
// memory and would emit flat instructions, which also count against the
// LDS wait counter, so that every LDS wait waits for them as well.  These
// name the global address space.
#define QOB_GLOBAL __attribute__((address_space(1)))
template <typename T> struct gvec;
template <> struct gvec<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct gvec<float2> { typedef float type __attribute__((ext_vector_type(2))); };
template <typename T> __device__ __forceinline__ T ldg(const T *p) {
    using V = typename gvec<T>::type;
    const V v = *(const QOB_GLOBAL V *)p;
    T r;
    r.x = v.x;
    r.y = v.y;
    return r;
}
// non-temporal load (data read once by this launch)
template <typename T> __device__ __forceinline__ T ldg_nt(const T *p) {
    using V = typename gvec<T>::type;
    const V v = __builtin_nontemporal_load((const QOB_GLOBAL V *)p);
    T r;
    r.x = v.x;
    r.y = v.y;
    return r;
}
template <typename T> __device__ __forceinline__ void stg(T *p, T v) {
    using V = typename gvec<T>::type;
    V w;
    w.x = v.x;
    w.y = v.y;
    *(QOB_GLOBAL V *)p = w;
}
template <typename T> __device__ __forceinline__ void stg_stream(T *p, T v) {
    using V = typename gvec<T>::type;
    V w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (QOB_GLOBAL V *)p);
}

// ------------------------------------------------------ fast integer divide
// q = n / d for 0 <= n < 2^31 with a host-computed multiplier:
//   l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
//   q = (umulhi(n, m) + n) >> l.
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, section 4.)  d == 2^l gives m == 1 and an
// exact shift.  Verified exhaustively against '/' on the host for the
// divisors used by the tests.
struct FastDiv {
    uint32_t d, m, l;
};
inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.l = l;
    f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
    return f;
}
__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t m, uint32_t l) {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t t = __umulhi(n, m);
#else
    uint32_t t = (uint32_t)(((uint64_t)n * m) >> 32);
#endif
    return (t + n) >> l;
}

// ---------------------------------------------------------- lazy gather plan
// A compiled sum of pure-sparse tensor-product terms (LazySum of
// LazyTensors whose factors are all sparse or identities), in gather form:
//   out[J] = alpha * sum_t sum_{entries of row jloc_t(J)} v * in[K] + beta * out[J]
// The term program is a flat int32 array walked with wave-uniform indices.
//
// Program layout (per term group):
//   [0] kind    : 0 = expanded local CSR, 1 = product of per-site tables
//   [1] flags   : bit0 = diagonal (all entries have K == J)
//   [2] q       : number of factor sites
//   [3] rowptr  : offset of the local row pointers in the rowptr pool (kind 0)
//                 or of the per-site tables in the entry pool (kind 1)
//   [4] fmask   : bitmask of factor sites (non-square shapes only)
//   then q site records of GP_SITE ints:
//     stride_l, m_stride, l_stride, d_l, m_d, l_d, lstride, tbl_off, stride_r
//   Digit of site s: j = (J / stride_l) mod d_l (l = output space,
//   r = input space; they differ only for non-square shapes).
//   kind 0: jloc = sum lstride * j; entries rowptr[jloc] .. rowptr[jloc+1]
//   kind 1: site table entry tbl_off + j holds (koff, v) with v == 0 meaning
//           an empty row.
enum { GP_HDR = 5, GP_SITE = 9 };

struct GatherPlanDev {
    const int32_t *prog = nullptr;
    int nterms = 0;
    const int32_t *rowptr = nullptr;   // local CSR row pointers
    const int32_t *koff = nullptr;     // per entry: delta (square) or r-offset
    const void *vals = nullptr;        // per entry value (dtype of the plan)
    // non-square support: all site digits are needed
    int nsites = 0;
    const int32_t *site_info = nullptr;  // per site: stride_l, m, l, d_l, m_d, l_d, stride_r, d_r
    bool square = true;
};

// ------------------------------------------------------------ launchers
// All launchers return hipSuccess or the launch error; they never allocate.

// elementwise (k_blas1.hip)
// Plan values from device-held coefficients (see k_refresh).
hipError_t launch_refresh(qob_dtype dt, int nslots, int nvals, const int32_t *ptr, const int32_t *term,
                          const double2 *bval, const int32_t *dst, const double2 *coef, void *vals, int32_t *prog,
                          hipStream_t s);
// Final stage of a two-stage FP64 reduction: *out = sum of nparts partials.
hipError_t launch_reduce_partials(int nparts, const double2 *partials, double2 *out, hipStream_t s);
// Compute units of the current device, queried once per device.
int device_cu_count();
hipError_t launch_scale(qob_dtype dt, int64_t n, double2 a, void *x, hipStream_t s);
hipError_t launch_copy(qob_dtype dt, int64_t n, const void *x, void *y, hipStream_t s);
hipError_t launch_axpy(qob_dtype dt, int64_t n, double2 a, const void *x, void *y, hipStream_t s);
hipError_t launch_axpby_copy(qob_dtype dt, int64_t n, double2 a, const void *x, double2 b, void *y, hipStream_t s);
hipError_t launch_embed(qob_dtype dt, const void *in, int64_t ri, int64_t ci, void *out, int64_t ro, int64_t co,
                        double2 a, double2 b, hipStream_t s);
hipError_t launch_convert(qob_dtype from, const void *src, qob_dtype to, void *dst, int64_t n, hipStream_t s);
// reductions into a device double2/double; partials must hold reduce_partials() doubles*2
int reduce_blocks(int64_t n);
hipError_t launch_dot(qob_dtype dt, int64_t n, const void *x, const void *y, double2 *partials, double2 *out, hipStream_t s);
hipError_t launch_nrm2(qob_dtype dt, int64_t n, const void *x, double2 *partials, double *out, hipStream_t s);
hipError_t launch_trace(qob_dtype dt, int64_t n, const void *A, double2 *partials, double2 *out, hipStream_t s);
// x *= 1/|norm| or 1/trace taken from a device scalar
hipError_t launch_scale_by_inv_norm(qob_dtype dt, int64_t n, const double *norm, void *x, hipStream_t s);
hipError_t launch_scale_by_inv_trace(qob_dtype dt, int64_t n, int64_t ld, const double2 *tr, void *A, hipStream_t s);

// sparse (k_sparse.hip)
// Sliced-ELL gather table of the rows of an m x n sparse matrix X:
// slices of 64 rows, slice k has width w_k; entry (p, r) of slice k sits at
// slice_off[k] + p*64 + r.  Column index -1 marks padding.
struct SellDev {
    int64_t m = 0, n = 0;
    int nslices = 0;
    int maxw = 0;                         // widest slice
    const int32_t *slice_off = nullptr;  // nslices+1 (in entries)
    const int32_t *col = nullptr;
    const void *val = nullptr;            // CF64 or CF32
    // Closed blocks (rows != nullptr): slice k holds rows rows[64k + lane]
    // (-1: idle lane); blocks of spb slices (4 or 5) hold whole connected
    // components of the pattern of X, and col[] holds the lane (0..64 spb - 1)
    // of the block that owns the needed row of B.  nslices is a multiple of
    // spb.
    const int32_t *rows = nullptr;
    int spb = 4;
    // Closed blocks whose lanes 2t and 2t + 1 hold rows 2i and 2i + 1 (or
    // are both idle): ComplexF32 then runs two rows per thread.
    bool pairs = false;
};
// CSC of X (columns of X as lists of (row, value)).
struct CscDev {
    int64_t m = 0, n = 0;
    const int32_t *colptr = nullptr, *row = nullptr;
    const void *val = nullptr;
    const int32_t *corder = nullptr;  // traversal order of the columns (nullptr: natural)
    int32_t sweep_bw = -1;  // largest distance in that order between columns sharing a row (-1: unknown)
    // Component groups (gint != nullptr): ngroups fixed-size records of
    // gmax (2, 4 or 8) columns: ints [nout, nin, mask_lo, mask_hi,
    // out[gmax], in[gmax]] and gmax x gmax coefficients (row = output).
    const int32_t *gint = nullptr;
    const void *gcoef = nullptr;
    int32_t ngroups = 0, gmax = 0;
};
// R (m x cols) = alpha * X * B + beta * R, B (n x cols)
hipError_t launch_sell_left(qob_dtype dt, const SellDev &X, const void *B, int64_t cols,
                            void *R, double2 alpha, double2 beta, hipStream_t s);
// R (rows x n) = alpha * B * X + beta * R, B (rows x m)
hipError_t launch_csc_right(qob_dtype dt, const CscDev &X, const void *B, int64_t rows,
                            void *R, double2 alpha, double2 beta, hipStream_t s);

// lazy gather (k_lazy.hip)
// out (n x nbatch, J fastest) = alpha * G in + beta * out, J = row index
hipError_t launch_gather_rows(qob_dtype dt, const GatherPlanDev &P, const void *in, int64_t n_in,
                              void *out, int64_t n_out, int64_t nbatch, double2 alpha,
                              double2 beta, hipStream_t s);
// out (nrows x n_out) = alpha * (G applied along the column index) + beta*out:
// out[:, J] = sum entries v * in[:, K]
hipError_t launch_gather_cols(qob_dtype dt, const GatherPlanDev &P, const void *in, int64_t n_in,
                              void *out, int64_t n_out, int64_t nrows, double2 alpha,
                              double2 beta, hipStream_t s);

// dense mode products (k_mode.hip)
// out viewed (n1, d_out, n3), in viewed (n1, d_in, n3), column-major;
// out = alpha * F x_2 in + beta * out, F is d_out x d_in column-major.
// ws: mode_workspace_bytes() bytes of device memory (null when that is 0).
hipError_t launch_mode(qob_dtype dt, const void *F, int64_t d_out, int64_t d_in, const void *in, void *out,
                       int64_t n1, int64_t n3, double2 alpha, double2 beta, hipStream_t s, void *ws);
size_t mode_workspace_bytes(int64_t d_out, int64_t d_in, int64_t n1, int64_t n3);
// y = alpha * F x + beta * y for an m x n column-major F (part: split
// partial sums, zgemv_splits(m, n) * m complex doubles when that is > 1), and
// y = alpha * A^T x + beta * y for a k x m column-major A.
int64_t zgemv_splits(int64_t m, int64_t n);
hipError_t launch_zgemv(qob_dtype dt, const void *F, int64_t m, int64_t n, const void *x, void *y, void *part,
                        double2 alpha, double2 beta, hipStream_t s);
hipError_t launch_zgemv_t(qob_dtype dt, const void *A, int64_t k, int64_t m, const void *x, void *y, double2 alpha,
                          double2 beta, hipStream_t s);
// C[m, n] = alpha * sum_k A[m, k] B[k, n] + beta * C[m, n] for z < batch, with
// element strides (A: m, k, batch; B: k, n, batch; C: m, n, batch); K > 0.
hipError_t launch_zgemm(qob_dtype dt, int64_t M, int64_t N, int64_t K, const void *A, int64_t sam, int64_t sak,
                        int64_t sa, const void *B, int64_t sbk, int64_t sbn, int64_t sb, void *C, int64_t scm,
                        int64_t scn, int64_t sc, int64_t batch, double2 alpha, double2 beta, hipStream_t s);

// fused dense LazyTensor on a column (k_fused.hip): modes applied in order,
// mode s acts on dimension d[s] with n1[s] = product of the dims before it;
// A[s] is the d x d factor (column-major, dtype of the launch).
enum { FUSED_MAXMODES = 8 };
struct FusedModes {
    int nmodes = 0;
    int D = 0;
    int n1[FUSED_MAXMODES] = {0};
    int d[FUSED_MAXMODES] = {0};
    const void *A[FUSED_MAXMODES] = {nullptr};
};
bool fused_supported(qob_dtype dt, const FusedModes &M);

// qubit Pauli-type LazySum (k_qubit.hip)
struct QubitStateDev {
    int32_t nbits = 0;          // state length 2^nbits
    int32_t prog_off = 0;       // program of the flips inside a tile (all flips if prog_hi_off < 0)
    int32_t prog_hi_off = -1;   // program of the flips of bits above the tile (second pass), or -1
    int32_t pad_ = 0;
    const void *in = nullptr;
    void *out = nullptr;
};
struct QubitItemDev {
    int32_t state;
    int32_t tile;
};
// Items split among the 8 XCDs of the device: the workgroups b with
// b % 8 == x (the ones the dispatcher places on XCD x) walk the items
// [off[x], off[x + 1]), so all tiles of one state share one L2.
struct QubitSplit {
    int32_t off[9];
};
// Batched form: items (state, tile) index a device table of states.  xs:
// per-XCD item ranges (the items grouped by XCD), or null.
hipError_t launch_qubit(qob_dtype dt, const QubitStateDev *states, const QubitItemDev *items, int nitems, int tbits,
                        const int32_t *prog, const void *vals, double2 alpha, double2 beta, double2 gamma,
                        hipStream_t s, const QubitSplit *xs = nullptr);
// Columns form: ncols states of one program, state c at in/out + c * 2^nbits.
hipError_t launch_qubit_cols(qob_dtype dt, QubitStateDev single, int64_t ncols, int tbits, const int32_t *prog,
                             const void *vals, double2 alpha, double2 beta, double2 gamma, hipStream_t s);
hipError_t launch_qubit_axpy(qob_dtype dt, const QubitStateDev *states, const QubitItemDev *items, int nitems,
                             int tbits, double2 a, hipStream_t s);
// Second pass over strided tiles (high-bit flips) for states with prog_hi_off >= 0:
// out = alpha * high + beta * out; the first pass then adds to out.
hipError_t launch_qubit_hi(qob_dtype dt, const QubitStateDev *states, const QubitItemDev *items, int nitems, int tbits,
                           const int32_t *prog, const void *vals, double2 alpha, double2 beta, hipStream_t s,
                           const QubitSplit *xs = nullptr);
hipError_t launch_qubit_hi_cols(qob_dtype dt, QubitStateDev single, int64_t ncols, int tbits, const int32_t *prog,
                                const void *vals, double2 alpha, double2 beta, hipStream_t s);
// out (D x ncols) = alpha * (A_n ... A_1) in + beta * out
hipError_t launch_lt_fused(qob_dtype dt, const FusedModes &M, const void *in, void *out, int64_t ncols,
                           double2 alpha, double2 beta, hipStream_t s, int64_t tile_ld = 0,
                           int64_t tile_nrb = 1);

// Factors on sites 1 and 2 (d1 = 16, d2 in {16, 32}) and optionally on the
// last site 3 (d3 in {16, 32}) of ncols columns: modes 1 and 2 chained in
// registers (k_ltreg.hip).  hipErrorNotSupported when the shape does not fit.
hipError_t launch_lt_reg(qob_dtype dt, const FusedModes &M, const void *in, void *out, int64_t ncols,
                         double2 alpha, double2 beta, hipStream_t s);
hipError_t launch_lt_reg_tile(qob_dtype dt, const FusedModes &M, const void *in, void *out, int64_t ntiles,
                              int64_t nrb, int64_t ld, double2 alpha, double2 beta, hipStream_t s);

// Fused reductions over a square gather plan (FP64 partials, one per block):
//   mode 0: sum_J conj(u[J]) (H v)[J]          (expect / bilinear on a Ket)
//   mode 1: sum_J (H V[:, J])[J], V n x n       (tr(H V), expect on rho)
int gather_dot_blocks(int64_t n);
hipError_t launch_gather_dot(qob_dtype dt, const GatherPlanDev &P, int mode, const void *u, const void *v, int64_t n,
                             double2 *partials, hipStream_t s);
// sum_J conj(u[J]) (H v)[J] for one qubit state (single.in = v), FP64
// partials, one per workgroup of the returned grid size
int qubit_dot_blocks(int64_t ntiles);
hipError_t launch_qubit_dot(qob_dtype dt, QubitStateDev single, const void *u, int tbits, const int32_t *prog,
                            const void *vals, double2 *partials, hipStream_t s);
// sum over the nonzeros of a CSC matrix X (n columns) of val * V[c + row * n]:
// tr(X V) for an n x n V (expect(::SparseOpPureType, ::Operator))
int csc_trace_blocks(int64_t n);
hipError_t launch_csc_trace(qob_dtype dt, const CscDev &X, const void *V, int64_t n, double2 *partials, hipStream_t s);

// ------------------------------------------------- host-side shared helpers
// Record the text of a failing call for qob_last_error() and return st.
qob_status set_error(qob_status st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// Device index of a context and the stream a call runs on (s, or the
// context's own stream when s is null).
int ctx_device(qob_ctx ctx);
hipStream_t ctx_stream(qob_ctx ctx, qob_stream s);
// Look up the no-allocation scratch array's address on `device` (current).
qob_status noalloc_init(int device);

}  // namespace qob
