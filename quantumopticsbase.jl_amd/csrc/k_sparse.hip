// Sparse x dense kernels: the SparseOperator branch of mul!.
//
// Left  R = alpha * X * B + beta * R   (src/sparsematrix.jl:25-47, 99-113;
//        gemv src/sparsematrix.jl:190-214; adjoint src/sparsematrix.jl:73-96,148-163)
// Right R = alpha * B * X + beta * R   (src/sparsematrix.jl:115-146;
//        vector x CSC src/sparsematrix.jl:216-238; adjoint src/sparsematrix.jl:49-71,165-175)
//
// The reference scatters (walks CSC columns and adds into result rows).  On
// the GPU both directions are written as gathers so that every output
// element is owned by one lane and written exactly once, fused with the
// beta scaling (the reference's separate fill!/rmul! pass):
//   left : lanes own rows of R and walk the rows of X, held as sliced ELL
//          (64-row slices, one wavefront per slice, coalesced index loads);
//          when the pattern of X splits into small connected components,
//          whole components share a workgroup and read B once through LDS;
//   right: lanes own rows of R (contiguous), one output column per step,
//          the column's CSC entries are wave-uniform (scalar loads) and each
//          gathers a contiguous run of a column of B; when the pattern splits
//          into components of at most 8 columns, one wave forms a whole
//          component from one load of each column of B it reads.
// The adjoint variants use the same kernels on the conjugate-transposed
// tables built at upload time.
#include "qob_internal.h"

#include <algorithm>

namespace qob {

namespace {

// Workgroups are dispatched round-robin over the 8 XCDs (linear id b runs on
// XCD b % 8).  xcd_remap turns b into a logical id such that each XCD owns a
// contiguous range of logical ids, so work that shares data shares an L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t G) {
    const uint32_t x = b & 7u, i = b >> 3;
    const uint32_t q = G >> 3, r = G & 7u;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

template <typename T> __device__ __forceinline__ T to_t(double2 a) {
    T r; r.x = (typename real_of<T>::type)a.x; r.y = (typename real_of<T>::type)a.y; return r;
}

// ---- left: sliced ELL row gather.  Block = 256 lanes = 4 slices of 64 rows.
// Logical block = (column group cg, slice group sg), slice group fastest so
// one XCD walks all row slices of the same columns of B.
template <typename T, int CJ>
__global__ void __launch_bounds__(256) k_sell_left(int nslices, int64_t m, const int32_t *__restrict__ slice_off,
                                                   const int32_t *__restrict__ col, const T *__restrict__ val,
                                                   const T *__restrict__ B, int64_t ldb, int64_t cols,
                                                   T *__restrict__ R, int64_t ldr, double2 alpha_, double2 beta_,
                                                   int beta_zero, uint32_t nsg) {
    const uint32_t G = gridDim.x;
    const uint32_t lb = xcd_remap(blockIdx.x, G);
    const uint32_t sg = lb % nsg, cg = lb / nsg;
    const int slice = (int)(sg * 4 + (threadIdx.x >> 6));
    const int r = threadIdx.x & 63;
    if (slice >= nslices) return;
    const int64_t row = (int64_t)slice * 64 + r;
    const int64_t j0 = (int64_t)cg * CJ;
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);

    T acc[CJ];
#pragma unroll
    for (int jj = 0; jj < CJ; ++jj) acc[jj] = czero<T>();

    const int32_t o0 = slice_off[slice];
    const int w = (slice_off[slice + 1] - o0) >> 6;
    for (int p = 0; p < w; ++p) {
        const int32_t c = col[o0 + p * 64 + r];
        if (c >= 0) {
            const T v = val[o0 + p * 64 + r];
#pragma unroll
            for (int jj = 0; jj < CJ; ++jj) {
                if (j0 + jj < cols) acc[jj] = cfma(v, B[(int64_t)c + (j0 + jj) * ldb], acc[jj]);
            }
        }
    }
    if (row < m) {
#pragma unroll
        for (int jj = 0; jj < CJ; ++jj) {
            if (j0 + jj < cols) {
                T *o = R + row + (j0 + jj) * ldr;
                st_out(o, axpby(alpha, acc[jj], beta, o, (beta_zero & 1) != 0), (beta_zero & 2) != 0);
            }
        }
    }
}

// ---- left, register-resident rows: each lane loads its row's (col, val)
// entries once and then walks a chunk of columns of B, UNR at a time, so the
// sparse tables are read once per workgroup instead of once per column
// group.  Logical block = (column chunk cg, slice group sg), slice group
// fastest.
template <typename T, int MAXW, int UNR>
__global__ void __launch_bounds__(256) k_sell_left_rr(int nslices, int64_t m, const int32_t *__restrict__ slice_off,
                                                      const int32_t *__restrict__ col, const T *__restrict__ val,
                                                      const T *__restrict__ B, int64_t ldb, int64_t cols,
                                                      T *__restrict__ R, int64_t ldr, double2 alpha_,
                                                      double2 beta_, int beta_zero, uint32_t nsg, int jchunk) {
    const uint32_t G = gridDim.x;
    const uint32_t lb = xcd_remap(blockIdx.x, G);
    const uint32_t sg = lb % nsg, cg = lb / nsg;
    const int slice = (int)(sg * 4 + (threadIdx.x >> 6));
    const int r = threadIdx.x & 63;
    if (slice >= nslices) return;
    const int64_t row = (int64_t)slice * 64 + r;
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    const int32_t o0 = slice_off[slice];
    const int w = (slice_off[slice + 1] - o0) >> 6;
    int32_t c[MAXW];
    T v[MAXW];
#pragma unroll
    for (int p = 0; p < MAXW; ++p) {
        c[p] = (p < w) ? col[o0 + p * 64 + r] : -1;
        v[p] = (p < w) ? val[o0 + p * 64 + r] : czero<T>();
    }
    const int64_t j0 = (int64_t)cg * jchunk;
    const int64_t j1 = (j0 + jchunk < cols) ? j0 + jchunk : cols;
    const bool live = row < m;
    for (int64_t j = j0; j < j1; j += UNR) {
        T acc[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) acc[u] = czero<T>();
#pragma unroll
        for (int p = 0; p < MAXW; ++p) {
            if (c[p] >= 0) {
#pragma unroll
                for (int u = 0; u < UNR; ++u)
                    if (j + u < j1) acc[u] = cfma(v[p], B[(int64_t)c[p] + (j + u) * ldb], acc[u]);
            }
        }
        if (live) {
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                if (j + u < j1) {
                    T *o = R + row + (j + u) * ldr;
                    st_out(o, axpby(alpha, acc[u], beta, o, (beta_zero & 1) != 0), (beta_zero & 2) != 0);
                }
            }
        }
    }
}

// ---- left, closed blocks: lane t of workgroup k owns row rows[BLK k + t]
// of R and of B (-1: idle), BLK = 64 SPB.  The block's rows are whole
// connected components of the pattern of X, so each entry of X in the block
// names a lane of the same block (col = that lane).  Per batch of UNR
// columns every lane loads its own row of B once into the workgroup's LDS
// and the entries read from there: each element of B is loaded once, with no
// re-reads through L2.
template <typename T, int MAXW, int UNR, int SPB>
__global__ void __launch_bounds__(64 * SPB) k_sell_left_closed(int nslices, const int32_t *__restrict__ slice_off,
                                                               const int32_t *__restrict__ rows,
                                                               const int32_t *__restrict__ col,
                                                               const T *__restrict__ val, const T *__restrict__ B,
                                                               int64_t ldb, int64_t cols, T *__restrict__ R,
                                                               int64_t ldr, double2 alpha_, double2 beta_,
                                                               int beta_zero, uint32_t nsg, int jchunk, int remap) {
    constexpr int BLK = 64 * SPB;
    __shared__ T sb[UNR][BLK];
    const uint32_t G = gridDim.x;
    const uint32_t lb = remap ? xcd_remap(blockIdx.x, G) : blockIdx.x;
    const uint32_t sg = lb % nsg, cg = lb / nsg;
    const int t = threadIdx.x;
    if ((int64_t)sg * SPB >= nslices) return;  // workgroup-uniform, before any barrier
    const int slice = (int)(sg * SPB + (t >> 6));  // nslices is a multiple of SPB: every lane has a slice
    const int r = t & 63;
    const int32_t row = rows[(int64_t)sg * BLK + t];
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    const int32_t o0 = slice_off[slice];
    const int w = (slice_off[slice + 1] - o0) >> 6;
    int32_t c[MAXW];
    T v[MAXW];
#pragma unroll
    for (int p = 0; p < MAXW; ++p) {
        c[p] = (p < w) ? col[o0 + p * 64 + r] : -1;
        v[p] = (p < w) ? val[o0 + p * 64 + r] : czero<T>();
    }
    const int64_t j0 = (int64_t)cg * jchunk;
    const int64_t j1 = (j0 + jchunk < cols) ? j0 + jchunk : cols;
    for (int64_t j = j0; j < j1; j += UNR) {
#pragma unroll
        for (int u = 0; u < UNR; ++u)
            sb[u][t] = (row >= 0 && j + u < j1) ? B[(int64_t)row + (j + u) * ldb] : czero<T>();
        __syncthreads();
        T acc[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) acc[u] = czero<T>();
#pragma unroll
        for (int p = 0; p < MAXW; ++p) {
            if (c[p] >= 0) {
#pragma unroll
                for (int u = 0; u < UNR; ++u) acc[u] = cfma(v[p], sb[u][c[p]], acc[u]);
            }
        }
        __syncthreads();  // every lane has read this batch before the next one is written
        if (row >= 0) {
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                if (j + u < j1) {
                    T *o = R + row + (j + u) * ldr;
                    st_out(o, axpby(alpha, acc[u], beta, o, (beta_zero & 1) != 0), (beta_zero & 2) != 0);
                }
            }
        }
    }
}

// ---- right: CSC column gather, lanes along rows of B/R.
// Column slot k of the sweep is output column corder[k] when a traversal
// order is given: an order that puts columns sharing a column of B next to
// each other makes the second read of that column an L2 hit.  Two sweeps:
//   colmajor = 0: logical block = (row block rb, column group cg), column
//     group fastest: each XCD sweeps a stripe of rows across all columns,
//     so B is reused within the stripe however far apart the columns are;
//   colmajor = 1: row block fastest: each XCD takes a contiguous range of
//     column slots and finishes every row block of a slot before the next,
//     so reads and writes stream whole columns.  B is reused in L2 only when
//     columns sharing a column of B are a few slots apart.
// With W > 1 a workgroup holds W waves of RB = 64 rows each; wave w takes
// the next cc column slots after wave w - 1 (a column group is W * cc slots).
template <typename T, int RB, int W = 1>
__global__ void __launch_bounds__(RB * W) k_csc_right(int64_t rows, int64_t n, const int32_t *__restrict__ colptr,
                                                      const int32_t *__restrict__ rowind, const T *__restrict__ val,
                                                      const T *__restrict__ B, T *__restrict__ R, double2 alpha_,
                                                      double2 beta_, int beta_zero, uint32_t ncg, int cc, int remap,
                                                      const int32_t *__restrict__ corder, int colmajor,
                                                      uint32_t nrb) {
    const uint32_t G = gridDim.x;
    const uint32_t lb = remap ? xcd_remap(blockIdx.x, G) : blockIdx.x;
    const uint32_t cg = colmajor ? lb / nrb : lb % ncg, rb = colmajor ? lb % nrb : lb / ncg;
    // wave index made explicitly wave-uniform so the column's CSC entries stay scalar loads
    const int wv = (W > 1) ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x / RB)) : 0;
    const int64_t i = (int64_t)rb * RB + (W > 1 ? threadIdx.x % RB : threadIdx.x);
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    const bool ok = i < rows;
    const int64_t c0 = ((int64_t)cg * W + wv) * cc;
    const int64_t c1 = (c0 + cc < n) ? c0 + cc : n;
    for (int64_t k = c0; k < c1; ++k) {
        const int64_t c = corder ? (int64_t)corder[k] : k;
        const int32_t p0 = colptr[c], p1 = colptr[c + 1];
        T acc = czero<T>();
        if (ok) {
            for (int32_t p = p0; p < p1; ++p) {
                acc = cfma(val[p], B[i + (int64_t)rowind[p] * rows], acc);
            }
            T *o = R + i + c * rows;
            st_out(o, axpby(alpha, acc, beta, o, (beta_zero & 1) != 0), (beta_zero & 2) != 0);
        }
    }
}

// ---- right, component groups: one wave per (group, 64-row block), row
// block fastest so each XCD streams a contiguous range of groups.  The
// group's record is wave-uniform: the columns of B it reads are loaded once
// into registers and every output column is formed from them, its stored
// entries (mask bits) in ascending input column.
template <typename T, int GM>
__global__ void __launch_bounds__(64) k_csc_right_groups(int64_t rows, const int32_t *__restrict__ gint,
                                                         const T *__restrict__ gcoef, const T *__restrict__ B,
                                                         T *__restrict__ R, double2 alpha_, double2 beta_,
                                                         int beta_zero, uint32_t nrb, int remap) {
    constexpr int NI = 4 + 2 * GM;
    const uint32_t G = gridDim.x;
    const uint32_t lb = remap ? xcd_remap(blockIdx.x, G) : blockIdx.x;
    const uint32_t g = lb / nrb, rb = lb % nrb;
    const int32_t *gi = gint + (int64_t)g * NI;
    const T *gc = gcoef + (int64_t)g * GM * GM;
    const int nout = gi[0], nin = gi[1];
    const uint64_t mask = (uint64_t)(uint32_t)gi[2] | ((uint64_t)(uint32_t)gi[3] << 32);
    const int64_t i = (int64_t)rb * 64 + threadIdx.x;
    const bool ok = i < rows;
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    T b[GM];
#pragma unroll
    for (int q = 0; q < GM; ++q) b[q] = (q < nin && ok) ? B[i + (int64_t)gi[4 + GM + q] * rows] : czero<T>();
#pragma unroll
    for (int o = 0; o < GM; ++o) {
        if (o < nout) {
            T acc = czero<T>();
#pragma unroll
            for (int q = 0; q < GM; ++q)
                if ((mask >> (o * GM + q)) & 1ull) acc = cfma(gc[o * GM + q], b[q], acc);
            if (ok) {
                T *out = R + i + (int64_t)gi[4 + o] * rows;
                st_out(out, axpby(alpha, acc, beta, out, (beta_zero & 1) != 0), (beta_zero & 2) != 0);
            }
        }
    }
}

typedef float f4v __attribute__((ext_vector_type(4)));

// The same for ComplexF32 with an even row count and 16-byte aligned B and
// R: each lane owns two adjacent rows, so every load and store moves 16
// bytes (a wave covers 128 rows).  Each row's arithmetic is the one above,
// in the same order.
template <int GM>
__global__ void __launch_bounds__(64) k_csc_right_groups_x2(int64_t rows, const int32_t *__restrict__ gint,
                                                            const float2 *__restrict__ gcoef,
                                                            const float2 *__restrict__ B, float2 *__restrict__ R,
                                                            double2 alpha_, double2 beta_, int beta_zero,
                                                            uint32_t nrb, int remap) {
    constexpr int NI = 4 + 2 * GM;
    const uint32_t G = gridDim.x;
    const uint32_t lb = remap ? xcd_remap(blockIdx.x, G) : blockIdx.x;
    const uint32_t g = lb / nrb, rb = lb % nrb;
    const int32_t *gi = gint + (int64_t)g * NI;
    const float2 *gc = gcoef + (int64_t)g * GM * GM;
    const int nout = gi[0], nin = gi[1];
    const uint64_t mask = (uint64_t)(uint32_t)gi[2] | ((uint64_t)(uint32_t)gi[3] << 32);
    const int64_t i = (int64_t)rb * 128 + 2 * threadIdx.x;  // rows i and i + 1 (rows is even)
    const bool ok = i < rows;
    const float2 alpha = to_t<float2>(alpha_), beta = to_t<float2>(beta_);
    float2 b0[GM], b1[GM];
#pragma unroll
    for (int q = 0; q < GM; ++q) {
        f4v v = {0.f, 0.f, 0.f, 0.f};
        if (q < nin && ok) v = *(const f4v *)(B + i + (int64_t)gi[4 + GM + q] * rows);
        b0[q] = make_float2(v.x, v.y);
        b1[q] = make_float2(v.z, v.w);
    }
#pragma unroll
    for (int o = 0; o < GM; ++o) {
        if (o < nout) {
            float2 a0 = czero<float2>(), a1 = czero<float2>();
#pragma unroll
            for (int q = 0; q < GM; ++q)
                if ((mask >> (o * GM + q)) & 1ull) {
                    a0 = cfma(gc[o * GM + q], b0[q], a0);
                    a1 = cfma(gc[o * GM + q], b1[q], a1);
                }
            if (ok) {
                float2 *out = R + i + (int64_t)gi[4 + o] * rows;
                const float2 r0 = axpby(alpha, a0, beta, out, (beta_zero & 1) != 0);
                const float2 r1 = axpby(alpha, a1, beta, out + 1, (beta_zero & 1) != 0);
                const f4v v = {r0.x, r0.y, r1.x, r1.y};
                if (beta_zero & 2) __builtin_nontemporal_store(v, (f4v *)out);
                else *(f4v *)out = v;
            }
        }
    }
}

// ---- left, closed blocks of row pairs (ComplexF32): lanes 2t and 2t + 1 of
// a block hold rows 2i and 2i + 1 (or are both idle), and thread t owns both,
// so each column of B and R moves through one 16-byte load and one 16-byte
// store per thread.  Slices, staging in LDS and the order in which a row's
// entries are summed are those of k_sell_left_closed.
template <int MAXW, int UNR, int SPB>
__global__ void __launch_bounds__(32 * SPB) k_sell_left_closed_x2(int nslices, const int32_t *__restrict__ slice_off,
                                                                  const int32_t *__restrict__ rows,
                                                                  const int32_t *__restrict__ col,
                                                                  const float2 *__restrict__ val,
                                                                  const float2 *__restrict__ B, int64_t ldb,
                                                                  int64_t cols, float2 *__restrict__ R, int64_t ldr,
                                                                  double2 alpha_, double2 beta_, int beta_zero,
                                                                  uint32_t nsg, int jchunk, int remap) {
    constexpr int BLK = 64 * SPB;
    __shared__ float2 sb[UNR][BLK];
    const uint32_t G = gridDim.x;
    const uint32_t lb = remap ? xcd_remap(blockIdx.x, G) : blockIdx.x;
    const uint32_t sg = lb % nsg, cg = lb / nsg;
    const int t = threadIdx.x;
    if ((int64_t)sg * SPB >= nslices) return;  // workgroup-uniform, before any barrier
    const int slice = (int)(sg * SPB + (t >> 5));
    const int r = 2 * (t & 31);  // first of the thread's two lanes in its slice
    const int32_t row = rows[(int64_t)sg * BLK + 2 * t];  // even, or -1 for an idle pair
    const float2 alpha = to_t<float2>(alpha_), beta = to_t<float2>(beta_);
    const int32_t o0 = slice_off[slice];
    const int w = (slice_off[slice + 1] - o0) >> 6;
    int2 c[MAXW];
    f4v v[MAXW];
#pragma unroll
    for (int p = 0; p < MAXW; ++p) {
        c[p] = (p < w) ? *(const int2 *)(col + o0 + p * 64 + r) : make_int2(-1, -1);
        v[p] = (p < w) ? *(const f4v *)(val + o0 + p * 64 + r) : f4v{0.f, 0.f, 0.f, 0.f};
    }
    const int64_t j0 = (int64_t)cg * jchunk;
    const int64_t j1 = (j0 + jchunk < cols) ? j0 + jchunk : cols;
    for (int64_t j = j0; j < j1; j += UNR) {
#pragma unroll
        for (int u = 0; u < UNR; ++u)
            *(f4v *)&sb[u][2 * t] = (row >= 0 && j + u < j1) ? *(const f4v *)(B + row + (j + u) * ldb)
                                                             : f4v{0.f, 0.f, 0.f, 0.f};
        __syncthreads();
        float2 a0[UNR], a1[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) a0[u] = a1[u] = czero<float2>();
#pragma unroll
        for (int p = 0; p < MAXW; ++p) {
            if (c[p].x >= 0) {
                const float2 x = make_float2(v[p].x, v[p].y);
#pragma unroll
                for (int u = 0; u < UNR; ++u) a0[u] = cfma(x, sb[u][c[p].x], a0[u]);
            }
            if (c[p].y >= 0) {
                const float2 x = make_float2(v[p].z, v[p].w);
#pragma unroll
                for (int u = 0; u < UNR; ++u) a1[u] = cfma(x, sb[u][c[p].y], a1[u]);
            }
        }
        __syncthreads();  // every lane has read this batch before the next one is written
        if (row >= 0) {
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                if (j + u < j1) {
                    float2 *o = R + row + (j + u) * ldr;
                    const float2 r0 = axpby(alpha, a0[u], beta, o, (beta_zero & 1) != 0);
                    const float2 r1 = axpby(alpha, a1[u], beta, o + 1, (beta_zero & 1) != 0);
                    const f4v q = {r0.x, r0.y, r1.x, r1.y};
                    if (beta_zero & 2) __builtin_nontemporal_store(q, (f4v *)o);
                    else *(f4v *)o = q;
                }
            }
        }
    }
}

// ---- right, short stripes: a wave covers RR rows x (64 / RR) columns per
// step (lane l: row l % RR, column l / RR).  Workgroups of one XCD sweep a
// stripe of RR rows across all columns, so a column of B re-read by the
// column delta positions later (the off-diagonals of X) is still in that
// XCD's L2: the stripe's B footprint is RR x n elements.
template <typename T, int RR>
__global__ void __launch_bounds__(64) k_csc_right_rr(int64_t rows, int64_t n, const int32_t *__restrict__ colptr,
                                                     const int32_t *__restrict__ rowind, const T *__restrict__ val,
                                                     const T *__restrict__ B, T *__restrict__ R, double2 alpha_,
                                                     double2 beta_, int beta_zero, uint32_t ncg, int cc, int remap) {
    constexpr int CW = 64 / RR;  // columns per step
    const uint32_t G = gridDim.x;
    const uint32_t lb = remap ? xcd_remap(blockIdx.x, G) : blockIdx.x;
    const uint32_t cg = lb % ncg, rb = lb / ncg;
    const int lane = threadIdx.x;
    const int64_t i = (int64_t)rb * RR + (lane % RR);
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    const bool ok = i < rows;
    const int64_t c0 = (int64_t)cg * cc * CW + lane / RR;
    for (int q = 0; q < cc; ++q) {
        const int64_t c = c0 + (int64_t)q * CW;
        if (!ok || c >= n) continue;
        const int32_t p0 = colptr[c], p1 = colptr[c + 1];
        T acc = czero<T>();
        for (int32_t p = p0; p < p1; ++p) acc = cfma(val[p], B[i + (int64_t)rowind[p] * rows], acc);
        T *o = R + i + c * rows;
        st_out(o, axpby(alpha, acc, beta, o, (beta_zero & 1) != 0), (beta_zero & 2) != 0);
    }
}

// ---- right with few rows (Bra, thin B): lanes along output columns.
template <typename T>
__global__ void __launch_bounds__(256) k_csc_right_thin(int64_t rows, int64_t n, const int32_t *__restrict__ colptr,
                                                        const int32_t *__restrict__ rowind, const T *__restrict__ val,
                                                        const T *__restrict__ B, T *__restrict__ R, double2 alpha_,
                                                        double2 beta_, int beta_zero) {
    const int64_t c = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (c >= n) return;
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    const int32_t p0 = colptr[c], p1 = colptr[c + 1];
    for (int64_t i = 0; i < rows; ++i) {
        T acc = czero<T>();
        for (int32_t p = p0; p < p1; ++p) acc = cfma(val[p], B[i + (int64_t)rowind[p] * rows], acc);
        T *o = R + i + c * rows;
        st_out(o, axpby(alpha, acc, beta, o, (beta_zero & 1) != 0), (beta_zero & 2) != 0);
    }
}

// tr(X V) = sum over nonzeros (row r, column c) of X of val * V[c, r]
// (expect(::SparseOpPureType, ::Operator), src/operators_sparse.jl:36-45):
// one pass over the nonzeros, one element of V each.
template <typename T>
__global__ void __launch_bounds__(256) k_csc_trace(int64_t n, const int32_t *__restrict__ colptr,
                                                   const int32_t *__restrict__ rowind, const T *__restrict__ val,
                                                   const T *__restrict__ V, double2 *__restrict__ partials) {
    __shared__ double2 sh[256];
    double2 r = make_double2(0, 0);
    for (int64_t c = blockIdx.x * (int64_t)256 + threadIdx.x; c < n; c += (int64_t)gridDim.x * 256) {
        for (int32_t p = colptr[c]; p < colptr[c + 1]; ++p) {
            const T a = val[p], b = V[c + (int64_t)rowind[p] * n];
            r.x += (double)a.x * b.x - (double)a.y * b.y;
            r.y += (double)a.x * b.y + (double)a.y * b.x;
        }
    }
    sh[threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[threadIdx.x].x += sh[threadIdx.x + w].x;
            sh[threadIdx.x].y += sh[threadIdx.x + w].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

}  // namespace

int csc_trace_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256)); }

hipError_t launch_csc_trace(qob_dtype dt, const CscDev &X, const void *V, int64_t n, double2 *partials, hipStream_t s) {
    const int g = csc_trace_blocks(n);
    if (dt == QOB_C64)
        hipLaunchKernelGGL(k_csc_trace<double2>, dim3(g), dim3(256), 0, s, n, X.colptr, X.row, (const double2 *)X.val,
                           (const double2 *)V, partials);
    else
        hipLaunchKernelGGL(k_csc_trace<float2>, dim3(g), dim3(256), 0, s, n, X.colptr, X.row, (const float2 *)X.val,
                           (const float2 *)V, partials);
    return hipGetLastError();
}

// Tunables (set once from the host; defaults chosen from measurements).
int g_sell_variant = 1;  // 0: column groups of CJ, 1: register-resident rows
int g_sell_cj = 4;
int g_sell_jchunk = 8;
int g_sell_unr = 4;  // columns in flight per lane in k_sell_left_rr (2, 4 or 8)
int g_csc_rb = 64;
int g_csc_cc = 2;
int g_csc_remap = 1;
int g_csc_rows = 64;  // 64: one column per wave step (k_csc_right); 16 or 32: short stripes
int g_csc_order = 1;  // 1: k_csc_right sweeps columns in the table's traversal order when it has one
int g_csc_waves = 1;  // waves per k_csc_right workgroup when csc_rb is 64 (1, 2 or 4)
int g_sell_cjchunk = 4;  // columns per workgroup of the closed-block left kernel
int g_csc_gkernel = 1;   // 1: the right product uses the table's component groups when it has them
// XCD remap of the group and closed-block kernels.  Both read each element
// of rho once, so there is no L2 reuse to keep on one XCD: the group kernel
// streams better with the plain round-robin dispatch (neighbouring blocks on
// different XCDs); the closed-block kernel measured faster remapped.
int g_csc_gremap = 0;
int g_sell_cremap = 1;
int g_csc_sweep = 2;  // k_csc_right sweep: 0 row stripes, 1 column-major, 2 column-major when sweep_bw <= 8
int g_nt_store = 1;

template <typename T>
static void sell_rr(const SellDev &X, const void *B, int64_t cols, void *R, double2 alpha, double2 beta, int bz,
                    hipStream_t s, int maxw, uint32_t nsg, uint32_t G, int jchunk) {
#define QOB_RR(MW, U)                                                                                          \
    hipLaunchKernelGGL((k_sell_left_rr<T, MW, U>), dim3(G), dim3(256), 0, s, X.nslices, X.m, X.slice_off, X.col, \
                       (const T *)X.val, (const T *)B, X.n, cols, (T *)R, X.m, alpha, beta, bz, nsg, jchunk)
    if (g_sell_unr == 8 && maxw <= 4) {
        if (maxw <= 2) QOB_RR(2, 8);
        else QOB_RR(4, 8);
    } else if (g_sell_unr == 2 && maxw <= 4) {
        if (maxw <= 2) QOB_RR(2, 2);
        else QOB_RR(4, 2);
    } else if (maxw <= 2) QOB_RR(2, 4);
    else if (maxw <= 4) QOB_RR(4, 4);
    else if (maxw <= 8) QOB_RR(8, 2);
    else QOB_RR(16, 1);
#undef QOB_RR
}

template <typename T>
static void sell_closed(const SellDev &X, const void *B, int64_t cols, void *R, double2 alpha, double2 beta, int bz,
                        hipStream_t s, uint32_t nsg, uint32_t G, int jchunk) {
#define QOB_CL(MW, U)                                                                                          \
    do {                                                                                                       \
        if (X.spb == 5)                                                                                        \
            hipLaunchKernelGGL((k_sell_left_closed<T, MW, U, 5>), dim3(G), dim3(320), 0, s, X.nslices,         \
                               X.slice_off, X.rows, X.col, (const T *)X.val, (const T *)B, X.n, cols, (T *)R,   \
                               X.m, alpha, beta, bz, nsg, jchunk, g_sell_cremap);                              \
        else                                                                                                   \
            hipLaunchKernelGGL((k_sell_left_closed<T, MW, U, 4>), dim3(G), dim3(256), 0, s, X.nslices,         \
                               X.slice_off, X.rows, X.col, (const T *)X.val, (const T *)B, X.n, cols, (T *)R,   \
                               X.m, alpha, beta, bz, nsg, jchunk, g_sell_cremap);                              \
    } while (0)
#define QOB_CL_U(MW)                         \
    do {                                     \
        if (jchunk == 1) QOB_CL(MW, 1);      \
        else if (jchunk == 2) QOB_CL(MW, 2); \
        else QOB_CL(MW, 4);                  \
    } while (0)
    if (X.maxw <= 2) QOB_CL_U(2);
    else if (X.maxw <= 4) QOB_CL_U(4);
    else if (X.maxw <= 8) QOB_CL_U(8);
    else QOB_CL_U(16);
#undef QOB_CL_U
#undef QOB_CL
}

static void sell_closed_x2(const SellDev &X, const void *B, int64_t cols, void *R, double2 alpha, double2 beta,
                           int bz, hipStream_t s, uint32_t nsg, uint32_t G, int jchunk) {
#define QOB_CL(MW, U)                                                                                          \
    do {                                                                                                       \
        if (X.spb == 5)                                                                                        \
            hipLaunchKernelGGL((k_sell_left_closed_x2<MW, U, 5>), dim3(G), dim3(160), 0, s, X.nslices,         \
                               X.slice_off, X.rows, X.col, (const float2 *)X.val, (const float2 *)B, X.n, cols, \
                               (float2 *)R, X.m, alpha, beta, bz, nsg, jchunk, g_sell_cremap);                 \
        else                                                                                                   \
            hipLaunchKernelGGL((k_sell_left_closed_x2<MW, U, 4>), dim3(G), dim3(128), 0, s, X.nslices,         \
                               X.slice_off, X.rows, X.col, (const float2 *)X.val, (const float2 *)B, X.n, cols, \
                               (float2 *)R, X.m, alpha, beta, bz, nsg, jchunk, g_sell_cremap);                 \
    } while (0)
#define QOB_CL_U(MW)                         \
    do {                                     \
        if (jchunk == 1) QOB_CL(MW, 1);      \
        else if (jchunk == 2) QOB_CL(MW, 2); \
        else QOB_CL(MW, 4);                  \
    } while (0)
    if (X.maxw <= 2) QOB_CL_U(2);
    else if (X.maxw <= 4) QOB_CL_U(4);
    else if (X.maxw <= 8) QOB_CL_U(8);
    else QOB_CL_U(16);
#undef QOB_CL_U
#undef QOB_CL
}

hipError_t launch_sell_left(qob_dtype dt, const SellDev &X, const void *B, int64_t cols, void *R,
                            double2 alpha, double2 beta, hipStream_t s) {
    if (X.m == 0 || cols == 0) return hipSuccess;
    const int bz = (beta.x == 0.0 && beta.y == 0.0) | (g_nt_store << 1);
    if (X.rows) {  // closed blocks (built only for square X, at most 16 entries per row)
        if (X.maxw > 16 || (X.spb != 4 && X.spb != 5) || X.nslices % X.spb) return hipErrorInvalidValue;
        const uint32_t nsg = (uint32_t)(X.nslices / X.spb);
        const int jchunk = (int)std::max<int64_t>(1, std::min<int64_t>(g_sell_cjchunk, cols));
        const uint64_t ncg = (uint64_t)((cols + jchunk - 1) / jchunk);
        const uint64_t G = (uint64_t)nsg * ncg;
        if (G > 0x7fffffffull) return hipErrorInvalidConfiguration;
        if (dt == QOB_C64) sell_closed<double2>(X, B, cols, R, alpha, beta, bz, s, nsg, (uint32_t)G, jchunk);
        else if (X.pairs && X.m % 2 == 0 && X.n % 2 == 0 && (((uintptr_t)B | (uintptr_t)R) & 15) == 0)
            sell_closed_x2(X, B, cols, R, alpha, beta, bz, s, nsg, (uint32_t)G, jchunk);
        else sell_closed<float2>(X, B, cols, R, alpha, beta, bz, s, nsg, (uint32_t)G, jchunk);
        return hipGetLastError();
    }
    if (g_sell_variant == 1 && X.maxw <= 16) {
        const uint32_t nsg = (uint32_t)((X.nslices + 3) / 4);
        const int jchunk = (int)std::max<int64_t>(1, std::min<int64_t>(g_sell_jchunk, cols));
        const uint64_t ncg = (uint64_t)((cols + jchunk - 1) / jchunk);
        const uint64_t G = (uint64_t)nsg * ncg;
        if (G > 0x7fffffffull) return hipErrorInvalidConfiguration;
        if (dt == QOB_C64) sell_rr<double2>(X, B, cols, R, alpha, beta, bz, s, X.maxw, nsg, (uint32_t)G, jchunk);
        else sell_rr<float2>(X, B, cols, R, alpha, beta, bz, s, X.maxw, nsg, (uint32_t)G, jchunk);
        return hipGetLastError();
    }
    const uint32_t nsg = (uint32_t)((X.nslices + 3) / 4);
    const int cj = (cols == 1) ? 1 : g_sell_cj;
    const uint32_t ncg = (uint32_t)((cols + cj - 1) / cj);
    const uint64_t G = (uint64_t)nsg * ncg;
    if (G > 0x7fffffffull) return hipErrorInvalidConfiguration;
#define QOB_SELL(TT, CJV)                                                                                 \
    hipLaunchKernelGGL((k_sell_left<TT, CJV>), dim3((uint32_t)G), dim3(256), 0, s, X.nslices, X.m,       \
                       X.slice_off, X.col, (const TT *)X.val, (const TT *)B, X.n, cols, (TT *)R, X.m, alpha, \
                       beta, bz, nsg)
    if (dt == QOB_C64) {
        if (cj == 1) QOB_SELL(double2, 1);
        else if (cj == 2) QOB_SELL(double2, 2);
        else if (cj == 4) QOB_SELL(double2, 4);
        else QOB_SELL(double2, 8);
    } else {
        if (cj == 1) QOB_SELL(float2, 1);
        else if (cj == 2) QOB_SELL(float2, 2);
        else if (cj == 4) QOB_SELL(float2, 4);
        else QOB_SELL(float2, 8);
    }
#undef QOB_SELL
    return hipGetLastError();
}

hipError_t launch_csc_right(qob_dtype dt, const CscDev &X, const void *B, int64_t rows, void *R,
                            double2 alpha, double2 beta, hipStream_t s) {
    if (X.n == 0 || rows == 0) return hipSuccess;
    const int bz = (beta.x == 0.0 && beta.y == 0.0) | (g_nt_store << 1);
    if (rows < 32) {
        const uint32_t g = (uint32_t)((X.n + 255) / 256);
        if (dt == QOB_C64)
            hipLaunchKernelGGL(k_csc_right_thin<double2>, dim3(g), dim3(256), 0, s, rows, X.n, X.colptr, X.row,
                               (const double2 *)X.val, (const double2 *)B, (double2 *)R, alpha, beta, bz);
        else
            hipLaunchKernelGGL(k_csc_right_thin<float2>, dim3(g), dim3(256), 0, s, rows, X.n, X.colptr, X.row,
                               (const float2 *)X.val, (const float2 *)B, (float2 *)R, alpha, beta, bz);
        return hipGetLastError();
    }
    if (X.gint && g_csc_gkernel) {
        const uint64_t nrb = (uint64_t)((rows + 63) / 64);
        const uint64_t G = nrb * (uint64_t)X.ngroups;
        if (G > 0x7fffffffull || X.ngroups <= 0) return hipErrorInvalidConfiguration;
#define QOB_CG(TT, GMV)                                                                                   \
    hipLaunchKernelGGL((k_csc_right_groups<TT, GMV>), dim3((uint32_t)G), dim3(64), 0, s, rows, X.gint,   \
                       (const TT *)X.gcoef, (const TT *)B, (TT *)R, alpha, beta, bz, (uint32_t)nrb, g_csc_gremap)
        if (dt == QOB_C64) {
            if (X.gmax == 2) QOB_CG(double2, 2);
            else if (X.gmax == 4) QOB_CG(double2, 4);
            else QOB_CG(double2, 8);
        } else if (rows % 2 == 0 && (((uintptr_t)B | (uintptr_t)R) & 15) == 0) {
            // two rows per lane: 128-row blocks
            const uint64_t nrb2 = (uint64_t)((rows + 127) / 128);
            const uint64_t G2 = nrb2 * (uint64_t)X.ngroups;
#define QOB_CG2(GMV)                                                                                           \
    hipLaunchKernelGGL((k_csc_right_groups_x2<GMV>), dim3((uint32_t)G2), dim3(64), 0, s, rows, X.gint,         \
                       (const float2 *)X.gcoef, (const float2 *)B, (float2 *)R, alpha, beta, bz, (uint32_t)nrb2, \
                       g_csc_gremap)
            if (X.gmax == 2) QOB_CG2(2);
            else if (X.gmax == 4) QOB_CG2(4);
            else QOB_CG2(8);
#undef QOB_CG2
        } else {
            if (X.gmax == 2) QOB_CG(float2, 2);
            else if (X.gmax == 4) QOB_CG(float2, 4);
            else QOB_CG(float2, 8);
        }
#undef QOB_CG
        return hipGetLastError();
    }
    if (g_csc_rows == 16 || g_csc_rows == 32) {
        const int rr = g_csc_rows, cw = 64 / rr, cc = g_csc_cc;
        const uint64_t nrb = (uint64_t)((rows + rr - 1) / rr);
        const uint64_t ncg = (uint64_t)((X.n + (int64_t)cc * cw - 1) / ((int64_t)cc * cw));
        const uint64_t G = nrb * ncg;
        if (G > 0x7fffffffull) return hipErrorInvalidConfiguration;
#define QOB_CRR(TT, RRV)                                                                                        \
    hipLaunchKernelGGL((k_csc_right_rr<TT, RRV>), dim3((uint32_t)G), dim3(64), 0, s, rows, X.n, X.colptr, X.row, \
                       (const TT *)X.val, (const TT *)B, (TT *)R, alpha, beta, bz, (uint32_t)ncg, cc, g_csc_remap)
        if (dt == QOB_C64) {
            if (rr == 16) QOB_CRR(double2, 16);
            else QOB_CRR(double2, 32);
        } else {
            if (rr == 16) QOB_CRR(float2, 16);
            else QOB_CRR(float2, 32);
        }
#undef QOB_CRR
        return hipGetLastError();
    }
    const int rb = g_csc_rb;
    const int cc = g_csc_cc;
    const int nw = (rb == 64 && (g_csc_waves == 2 || g_csc_waves == 4)) ? g_csc_waves : 1;
    const uint32_t nrb = (uint32_t)((rows + rb - 1) / rb);
    const uint32_t ncg = (uint32_t)((X.n + (int64_t)cc * nw - 1) / ((int64_t)cc * nw));
    const uint64_t G = (uint64_t)nrb * ncg;
    if (G > 0x7fffffffull) return hipErrorInvalidConfiguration;
    const int32_t *corder = g_csc_order ? X.corder : nullptr;
    // the column-major sweep needs the reuse distance of the order swept
    const int64_t bw = corder ? X.sweep_bw : (X.corder ? -1 : X.sweep_bw);
    const int colmajor = g_csc_sweep == 1 || (g_csc_sweep == 2 && bw >= 0 && bw <= 8);
#define QOB_CSC(TT, RBV, WV)                                                                                     \
    hipLaunchKernelGGL((k_csc_right<TT, RBV, WV>), dim3((uint32_t)G), dim3(RBV * WV), 0, s, rows, X.n, X.colptr, \
                       X.row, (const TT *)X.val, (const TT *)B, (TT *)R, alpha, beta, bz, ncg, cc, g_csc_remap,  \
                       corder, colmajor, nrb)
    if (dt == QOB_C64) {
        if (rb == 64) {
            if (nw == 4) QOB_CSC(double2, 64, 4);
            else if (nw == 2) QOB_CSC(double2, 64, 2);
            else QOB_CSC(double2, 64, 1);
        } else if (rb == 128) QOB_CSC(double2, 128, 1);
        else QOB_CSC(double2, 256, 1);
    } else {
        if (rb == 64) {
            if (nw == 4) QOB_CSC(float2, 64, 4);
            else if (nw == 2) QOB_CSC(float2, 64, 2);
            else QOB_CSC(float2, 64, 1);
        } else if (rb == 128) QOB_CSC(float2, 128, 1);
        else QOB_CSC(float2, 256, 1);
    }
#undef QOB_CSC
    return hipGetLastError();
}

}  // namespace qob
