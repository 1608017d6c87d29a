// Dense mode products and dense operator products on MFMA.
//
// Reference: the mode product of a tensor-product factor, _tp_matmul!
// (src/operators_lazytensor.jl:406-428), which dispatches to
// _tp_matmul_first! (:281-290), _tp_matmul_last! (:292-301) or
// _tp_matmul_mid! (:333-404, permutes through cached temporaries), and the
// DenseOperator mul! through BLAS gemm / gemv (src/operators_dense.jl:394-396).
// All of them are one operation here:
//   out[i1, a, i3] = alpha * sum_b F[a, b] in[i1, b, i3] + beta * out[i1, a, i3]
// with in viewed (n1, d_in, n3) and out (n1, d_out, n3), column-major, and F
// the d_out x d_in factor (column-major; a transposed or conjugated variant
// is materialised by the caller).  No permutation is ever materialised.
//
// Kernels (every complex product from four real MFMAs, see qob_mfma.h):
//   k_mode_fib    n1 >= 16, d_in, d_out <= 64.  A wave takes 16 fibres with
//                 consecutive i1 (one i3): their elements are read from HBM
//                 straight into MFMA B-fragments (16 consecutive i1 = 256 B
//                 runs), the factor's A-fragments come from LDS (staged once
//                 per workgroup), and the D tiles go straight back to HBM
//                 (again 256 B runs).  Each element is read and written once.
//   k_mode_chunk  n1 < 16 (the factor on the first, fastest site, or sites
//                 of small dimension before it), d_in, d_out <= 64.  A chunk
//                 of consecutive i3 blocks is contiguous in memory: it is
//                 loaded linearly into LDS planes, the mode product runs from
//                 LDS into output planes, and the output chunk is written
//                 back linearly.
//   k_zgemm       everything wider: a 64 x 64 x 16 LDS-tiled complex GEMM
//                 with arbitrary element strides and a batch index, used for
//                 factors wider than 64 and for DenseOperator on a density
//                 matrix (and qob_gemm).
//   k_zgemv       a single vector (DenseOperator on a Ket or Bra): columns
//                 of F split over workgroups, deterministic two-stage sum.
#include "qob_mfma.h"

#include <map>
#include <mutex>
#include <numeric>
#include <unordered_set>

namespace qob {

namespace {

__device__ __forceinline__ int pad16(int i) { return i + (i >> 4); }

constexpr int MODE_NT = 512;  // threads per workgroup of the mode kernels

// F (d_out x d_in, column-major) into Fs[i + k * ldf] for i < rows,
// k < cols, zero outside F.
template <typename T>
__device__ void stage_factor(T *Fs, int ldf, int rows, int cols, const T *__restrict__ F, int d_out, int d_in) {
    for (int e = threadIdx.x; e < rows * cols; e += blockDim.x) {
        const int i = e % rows, k = e / rows;
        Fs[i + k * ldf] = (i < d_out && k < d_in) ? F[i + (int64_t)k * d_out] : czero<T>();
    }
}

// ------------------------------------------------------------ k_mode_fib
// NT16 = ceil(max(d_in, d_out) / 16): output tiles per fibre group and
// k-steps / 4.  A fibre group is 16 fibres: with n1 >= 16, 16 consecutive
// i1 of one i3 (gpi groups per i3); with n1 < 16 (flat), 16 consecutive
// fibres f = i1 + n1 * i3 of the nf = n1 * n3 fibres.  The loads of the next
// chunk of 16 k -- the next group's first chunk after a group's last -- are
// in flight during the current chunk's MFMAs and the group's stores.
struct FibreAt {
    const void *p;  // first element of the fibre (in), or null past the end
    int64_t o;      // offset of the fibre in out
};
template <typename R, int NT16>
__global__ void __launch_bounds__(MODE_NT) k_mode_fib(const cplx<R> *__restrict__ F, int d_out, int d_in,
                                                       const cplx<R> *__restrict__ in, cplx<R> *__restrict__ out,
                                                       int n1, int gpi, int64_t ngroups, int64_t nf,
                                                       double2 alpha_, double2 beta_, int beta_zero, int flat) {
    using T = cplx<R>;
    using V = typename mf<R>::v4;
    constexpr int KS = 4 * NT16, LDF = 16 * NT16 + 1;
    extern __shared__ unsigned char lds_raw[];
    T *Fs = (T *)lds_raw;
    stage_factor(Fs, LDF, 16 * NT16, 4 * KS, F, d_out, d_in);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const T alpha = mk((R)alpha_.x, (R)alpha_.y), beta = mk((R)beta_.x, (R)beta_.y);
    const int64_t sin = (int64_t)n1 * d_in, sout = (int64_t)n1 * d_out;
    constexpr int WPB = MODE_NT / 64;
    const int64_t gstep = (int64_t)gridDim.x * WPB;
    // this lane's fibre of group g
    auto fibre = [&](int64_t g) -> FibreAt {
        int64_t i3;
        int i1;
        bool ok;
        if (g >= ngroups) return {nullptr, 0};
        if (flat) {
            const int64_t f = g * 16 + lr;
            ok = f < nf;
            i3 = n1 == 1 ? f : f / n1;
            i1 = (int)(f - i3 * n1);
        } else {
            i3 = g / gpi;
            i1 = (int)(g - i3 * gpi) * 16 + lr;
            ok = i1 < n1;
        }
        if (!ok) return {nullptr, 0};
        return {in + i1 + i3 * sin, i1 + i3 * sout};
    };
    // Every lane issues all of its loads (from an in-bounds address where
    // the element does not exist, followed by a zero select): with loads
    // under a branch the compiler must assume that none was issued after
    // the chunk in use, and waited for the next chunk's loads before this
    // chunk's MFMAs.  (16, 32, 16) chain on an 8192^2 rho: 1.211 -> 1.185 ms.
    auto load = [&](const FibreAt &fa, int kc, T (&x)[4]) {
        const T *base = fa.p ? (const T *)fa.p : in;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int b = 16 * kc + 4 * j + lk;
            const bool ok = fa.p && b < d_in;
            const T v = ldg(base + (int64_t)n1 * (ok ? b : 0));
            x[j] = ok ? v : czero<T>();
        }
    };
    int64_t g = (int64_t)blockIdx.x * WPB + wave;
    FibreAt cur = fibre(g);
    T x[4], xn[4];
    load(cur, 0, x);
    for (; g < ngroups; g += gstep) {
        const FibreAt nxt = fibre(g + gstep);
        V ar[NT16], ai[NT16];
#pragma unroll
        for (int t = 0; t < NT16; ++t) {
            ar[t] = V{0, 0, 0, 0};
            ai[t] = V{0, 0, 0, 0};
        }
#pragma unroll 1
        for (int kc = 0; kc < NT16; ++kc) {
            if (kc + 1 < NT16) load(cur, kc + 1, xn);
            else load(nxt, 0, xn);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
#pragma unroll
                for (int t = 0; t < NT16; ++t)
                    cmma4<R>(Fs[(16 * t + lr) + (16 * kc + 4 * j + lk) * LDF], x[j], ar[t], ai[t]);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) x[j] = xn[j];
        }
        if (cur.p) {
            T *ob = out + cur.o;
#pragma unroll
            for (int t = 0; t < NT16; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = 16 * t + mf<R>::row(lane, r);
                    if (a < d_out) {
                        T *o = ob + (int64_t)n1 * a;
                        T v = cmul(alpha, mk((R)ar[t][r], (R)ai[t][r]));
                        if (!beta_zero) v = cfma(beta, *o, v);
                        st_stream(o, v);
                    }
                }
            }
        }
        cur = nxt;
    }
}

// ---------------------------------------------------------- k_mode_chunk
// Chunks of cpc consecutive i3 blocks (cpc * n1 * d elements, contiguous).
constexpr int CHUNK_NT = 256;  // threads per workgroup: small chunks, several workgroups per CU
// input elements per thread loaded one chunk ahead (a ComplexF64 chunk of
// 32 KB of planes holds 512 input elements for square factors, ComplexF32 1024)
template <typename R> constexpr int chunk_pf() { return sizeof(R) == 8 ? 2 : 4; }
template <typename R, int NT16>
__global__ void __launch_bounds__(CHUNK_NT) k_mode_chunk(const cplx<R> *__restrict__ F, int d_out, int d_in,
                                                         const cplx<R> *__restrict__ in, cplx<R> *__restrict__ out,
                                                         int n1, int64_t n3, int cpc, int64_t nchunks,
                                                         double2 alpha_, double2 beta_, int beta_zero) {
    using T = cplx<R>;
    using V = typename mf<R>::v4;
    constexpr int KS = 4 * NT16, LDF = 16 * NT16 + 1;
    extern __shared__ unsigned char lds_raw[];
    T *Fs = (T *)lds_raw;
    const int pin = pad16(cpc * n1 * d_in) + 1, pout = pad16(cpc * n1 * d_out) + 1;
    R *ire = (R *)(Fs + 4 * KS * LDF);
    R *iim = ire + pin;
    R *ore = iim + pin;
    R *oim = ore + pout;
    stage_factor(Fs, LDF, 16 * NT16, 4 * KS, F, d_out, d_in);  // ordered by the first barrier below
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const T alpha = mk((R)alpha_.x, (R)alpha_.y), beta = mk((R)beta_.x, (R)beta_.y);
    // The first CHUNK_PF * CHUNK_NT input elements of the next chunk are
    // loaded into registers while this chunk is multiplied and written out
    // (the rest of a larger chunk is loaded when it is staged).
    auto chunk_in = [&](int64_t c, int &ne) -> const T * {
        const int64_t i30 = c * cpc;
        const int nb = (int)(n3 - i30 < cpc ? n3 - i30 : cpc);
        ne = nb * n1 * d_in;
        return in + i30 * n1 * d_in;
    };
    constexpr int CHUNK_PF = chunk_pf<R>();
    T pre[CHUNK_PF];
    auto prefetch = [&](int64_t c) {
        int ne = 0;
        const T *src = c < nchunks ? chunk_in(c, ne) : in;
#pragma unroll
        for (int j = 0; j < CHUNK_PF; ++j) {
            const int e = tid + j * CHUNK_NT;
            pre[j] = e < ne ? ldg(src + e) : czero<T>();
        }
    };
    prefetch(blockIdx.x);
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int64_t i30 = c * cpc;
        const int nb = (int)(n3 - i30 < cpc ? n3 - i30 : cpc);
        const int ne_out = nb * n1 * d_out;
        int ne_in;
        const T *src = chunk_in(c, ne_in);
#pragma unroll
        for (int j = 0; j < CHUNK_PF; ++j) {
            const int e = tid + j * CHUNK_NT;
            if (e < ne_in) {
                ire[pad16(e)] = pre[j].x;
                iim[pad16(e)] = pre[j].y;
            }
        }
        for (int e = tid + CHUNK_PF * CHUNK_NT; e < ne_in; e += CHUNK_NT) {
            const T v = ldg(src + e);
            ire[pad16(e)] = v.x;
            iim[pad16(e)] = v.y;
        }
        prefetch(c + gridDim.x);
        // input planes loaded; the previous chunk's output planes are copied out
        __syncthreads();
        const int nf = nb * n1, ng = (nf + 15) >> 4;
        for (int g = wave; g < ng; g += CHUNK_NT / 64) {
            const int f = g * 16 + lr;
            const bool ok = f < nf;
            const int i3l = ok ? f / n1 : 0, i1 = ok ? f - i3l * n1 : 0;
            const int bin = i1 + n1 * d_in * i3l;
            V ar[NT16], ai[NT16];
#pragma unroll
            for (int t = 0; t < NT16; ++t) {
                ar[t] = V{0, 0, 0, 0};
                ai[t] = V{0, 0, 0, 0};
            }
#pragma unroll 1
            for (int kc = 0; kc < NT16; ++kc) {
                T x[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int b = 16 * kc + 4 * j + lk;
                    if (ok && b < d_in) {
                        const int idx = pad16(bin + n1 * b);
                        x[j] = mk(ire[idx], iim[idx]);
                    } else {
                        x[j] = czero<T>();
                    }
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
#pragma unroll
                    for (int t = 0; t < NT16; ++t)
                        cmma4<R>(Fs[(16 * t + lr) + (16 * kc + 4 * j + lk) * LDF], x[j], ar[t], ai[t]);
                }
            }
            if (ok) {
                const int bout = i1 + n1 * d_out * i3l;
#pragma unroll
                for (int t = 0; t < NT16; ++t) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int a = 16 * t + mf<R>::row(lane, r);
                        if (a < d_out) {
                            const int idx = pad16(bout + n1 * a);
                            ore[idx] = ar[t][r];
                            oim[idx] = ai[t][r];
                        }
                    }
                }
            }
        }
        // output planes complete; every wave is done reading the input planes
        __syncthreads();
        T *dst = out + i30 * n1 * d_out;
        for (int e = tid; e < ne_out; e += CHUNK_NT) {
            T v = cmul(alpha, mk(ore[pad16(e)], oim[pad16(e)]));
            if (!beta_zero) v = cfma(beta, dst[e], v);
            st_stream(dst + e, v);
        }
    }
}

// --------------------------------------------------------------- k_zgemm
// C[m, n] = alpha * sum_k A[m, k] B[k, n] + beta * C[m, n] for every batch
// index z, with element (m, k) of A at A + z sa + m sam + k sak, (k, n) of B
// at B + z sb + k sbk + n sbn, (m, n) of C at C + z sc + m scm + n scn.
// Tiles of BM x BN outputs, K in chunks of BK through double-buffered LDS
// planes (real and imaginary parts apart).  Four waves in a 2 x 2 grid; the
// MFMA rows run along n and its columns (lane % 16) along m, so a C column
// with scm = 1 is stored as 16-element runs.
constexpr int GBM = 64, GBN = 64, GBK = 16, GPAD = 4, GNT = 256;

template <typename R>
__global__ void __launch_bounds__(GNT) k_zgemm(int64_t M, int64_t N, int64_t K, const cplx<R> *__restrict__ A,
                                               int64_t sam, int64_t sak, int64_t sa, const cplx<R> *__restrict__ B,
                                               int64_t sbk, int64_t sbn, int64_t sb, cplx<R> *__restrict__ C,
                                               int64_t scm, int64_t scn, int64_t sc, int64_t tiles_m, int64_t tiles_n,
                                               double2 alpha_, double2 beta_, int beta_zero) {
    using T = cplx<R>;
    using V = typename mf<R>::v4;
    constexpr int LA = GBM + GPAD, LB = GBN + GPAD;
    constexpr int AE = GBM * GBK / GNT, BE = GBN * GBK / GNT;
    constexpr int TMW = GBM / 32, TNW = GBN / 32;  // 16 x 16 tiles per wave along m and n
    extern __shared__ unsigned char lds_raw[];
    R *As = (R *)lds_raw;             // [buf][plane][k][m]
    R *Bs = As + 2 * 2 * GBK * LA;    // [buf][plane][k][n]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int wm = wave & 1, wn = wave >> 1;
    int64_t bid = blockIdx.x;
    const int64_t tn = bid % tiles_n;
    bid /= tiles_n;
    const int64_t tm = bid % tiles_m, z = bid / tiles_m;
    A += z * sa;
    B += z * sb;
    C += z * sc;
    const int64_t m0 = tm * GBM, n0 = tn * GBN;
    const bool amf = sam <= sak;  // consecutive threads along m (else along k)
    const bool bnf = sbn <= sbk;  // consecutive threads along n (else along k)
    T ra[AE], rb[BE];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < AE; ++i) {
            const int e = tid + GNT * i;
            const int m = amf ? e % GBM : e / GBK, k = amf ? e / GBM : e % GBK;
            ra[i] = (m0 + m < M && k0 + k < K) ? A[(m0 + m) * sam + (k0 + k) * sak] : czero<T>();
        }
#pragma unroll
        for (int i = 0; i < BE; ++i) {
            const int e = tid + GNT * i;
            const int n = bnf ? e % GBN : e / GBK, k = bnf ? e / GBN : e % GBK;
            rb[i] = (n0 + n < N && k0 + k < K) ? B[(k0 + k) * sbk + (n0 + n) * sbn] : czero<T>();
        }
    };
    auto store = [&](int buf) {
        R *a = As + buf * 2 * GBK * LA, *b = Bs + buf * 2 * GBK * LB;
#pragma unroll
        for (int i = 0; i < AE; ++i) {
            const int e = tid + GNT * i;
            const int m = amf ? e % GBM : e / GBK, k = amf ? e / GBM : e % GBK;
            a[k * LA + m] = ra[i].x;
            a[GBK * LA + k * LA + m] = ra[i].y;
        }
#pragma unroll
        for (int i = 0; i < BE; ++i) {
            const int e = tid + GNT * i;
            const int n = bnf ? e % GBN : e / GBK, k = bnf ? e / GBN : e % GBK;
            b[k * LB + n] = rb[i].x;
            b[GBK * LB + k * LB + n] = rb[i].y;
        }
    };
    V cr[TNW][TMW], ci[TNW][TMW];
#pragma unroll
    for (int p = 0; p < TNW; ++p)
#pragma unroll
        for (int q = 0; q < TMW; ++q) {
            cr[p][q] = V{0, 0, 0, 0};
            ci[p][q] = V{0, 0, 0, 0};
        }
    const int64_t nk = (K + GBK - 1) / GBK;
    load(0);
    store(0);
    __syncthreads();
    for (int64_t kc = 0; kc < nk; ++kc) {
        const int buf = (int)(kc & 1);
        if (kc + 1 < nk) load((kc + 1) * GBK);  // in flight during this chunk's MFMAs
        const R *a = As + buf * 2 * GBK * LA, *b = Bs + buf * 2 * GBK * LB;
#pragma unroll
        for (int ks = 0; ks < GBK / 4; ++ks) {
            const int k = 4 * ks + lk;
            T af[TNW], bf[TMW];
#pragma unroll
            for (int p = 0; p < TNW; ++p) {
                const int n = wn * (GBN / 2) + 16 * p + lr;
                af[p] = mk(b[k * LB + n], b[GBK * LB + k * LB + n]);
            }
#pragma unroll
            for (int q = 0; q < TMW; ++q) {
                const int m = wm * (GBM / 2) + 16 * q + lr;
                bf[q] = mk(a[k * LA + m], a[GBK * LA + k * LA + m]);
            }
#pragma unroll
            for (int p = 0; p < TNW; ++p)
#pragma unroll
                for (int q = 0; q < TMW; ++q) cmma4<R>(af[p], bf[q], cr[p][q], ci[p][q]);
        }
        if (kc + 1 < nk) store(buf ^ 1);  // buf ^ 1 was last read before the previous barrier
        __syncthreads();
    }
    const T alpha = mk((R)alpha_.x, (R)alpha_.y), beta = mk((R)beta_.x, (R)beta_.y);
#pragma unroll
    for (int p = 0; p < TNW; ++p)
#pragma unroll
        for (int q = 0; q < TMW; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t n = n0 + wn * (GBN / 2) + 16 * p + mf<R>::row(lane, r);
                const int64_t m = m0 + wm * (GBM / 2) + 16 * q + lr;
                if (m < M && n < N) {
                    T *c = C + m * scm + n * scn;
                    T v = cmul(alpha, mk((R)cr[p][q][r], (R)ci[p][q][r]));
                    if (!beta_zero) v = cfma(beta, *c, v);
                    *c = v;
                }
            }
}

// --------------------------------------------------------------- k_zgemv
// part[s][i] = sum over the columns j of split s of F[i, j] x[j]; with one
// split the result goes straight to y (y = alpha part + beta y).  A
// workgroup covers 64 rows (one row per lane, so every column of F is read as
// a 1 KiB run); its 4 waves take every 4th column of the split; the wave
// partials are summed in LDS in wave order.
constexpr int GV_NT = 256;
template <typename R>
__global__ void __launch_bounds__(GV_NT) k_zgemv(const cplx<R> *__restrict__ F, int64_t m, int64_t n,
                                                 const cplx<R> *__restrict__ x, int64_t split_len,
                                                 cplx<double> *__restrict__ part, cplx<R> *__restrict__ y,
                                                 double2 alpha_, double2 beta_, int beta_zero) {
    using T = cplx<R>;
    __shared__ T red[GV_NT];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const int64_t j0 = (int64_t)blockIdx.y * split_len;
    const int64_t j1 = j0 + split_len < n ? j0 + split_len : n;
    T acc = czero<T>();
    if (i < m) {
        const T *Fi = F + i;
        int64_t j = j0 + wave;
        for (; j + 12 < j1; j += 16) {
            const T f0 = Fi[j * m], f1 = Fi[(j + 4) * m], f2 = Fi[(j + 8) * m], f3 = Fi[(j + 12) * m];
            acc = cfma(f0, x[j], acc);
            acc = cfma(f1, x[j + 4], acc);
            acc = cfma(f2, x[j + 8], acc);
            acc = cfma(f3, x[j + 12], acc);
        }
        for (; j < j1; j += 4) acc = cfma(Fi[j * m], x[j], acc);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (wave == 0 && i < m) {
        T s = red[lane];
        s = cadd(s, red[64 + lane]);
        s = cadd(s, red[128 + lane]);
        s = cadd(s, red[192 + lane]);
        if (gridDim.y == 1) {
            const T alpha = mk((R)alpha_.x, (R)alpha_.y), beta = mk((R)beta_.x, (R)beta_.y);
            y[i] = axpby(alpha, s, beta, y + i, beta_zero != 0);
        } else {
            part[(int64_t)blockIdx.y * m + i] = mk((double)s.x, (double)s.y);
        }
    }
}

// y[i] = alpha * sum_s part[s][i] + beta * y[i], splits summed in order
template <typename R>
__global__ void __launch_bounds__(256) k_zgemv_sum(const cplx<double> *__restrict__ part, int nsplit, int64_t m,
                                                   cplx<R> *__restrict__ y, double2 alpha_, double2 beta_,
                                                   int beta_zero) {
    using T = cplx<R>;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    double2 s = part[i];
    for (int k = 1; k < nsplit; ++k) s = cadd(s, part[(int64_t)k * m + i]);
    const T alpha = mk((R)alpha_.x, (R)alpha_.y), beta = mk((R)beta_.x, (R)beta_.y);
    y[i] = axpby(alpha, mk((R)s.x, (R)s.y), beta, y + i, beta_zero != 0);
}

// y[i] = alpha * sum_l A[l + i k] x[l] + beta * y[i] (y = A^T x for a
// column-major k x m A): one wave per output, lanes along the column (1 KiB
// runs), then a butterfly sum over the wave in a fixed order.
template <typename R>
__global__ void __launch_bounds__(256) k_zgemv_t(const cplx<R> *__restrict__ A, int64_t k, int64_t m,
                                                 const cplx<R> *__restrict__ x, cplx<R> *__restrict__ y,
                                                 double2 alpha_, double2 beta_, int beta_zero) {
    using T = cplx<R>;
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= m) return;
    const T *Ai = A + i * k;
    T acc = czero<T>();
    for (int64_t l = lane; l < k; l += 64) acc = cfma(Ai[l], x[l], acc);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        acc.x += __shfl_xor(acc.x, o);
        acc.y += __shfl_xor(acc.y, o);
    }
    if (lane == 0) {
        const T alpha = mk((R)alpha_.x, (R)alpha_.y), beta = mk((R)beta_.x, (R)beta_.y);
        y[i] = axpby(alpha, acc, beta, y + i, beta_zero != 0);
    }
}

// Workgroups of a kernel resident per CU at the given block size and LDS
// (the runtime's occupancy query, cached per kernel and LDS size).
template <typename K> int resident_per_cu(K kfn, int nt, size_t lds) {
    static std::mutex mu;
    static std::map<std::pair<const void *, size_t>, int> cache;
    std::lock_guard<std::mutex> g(mu);
    auto key = std::make_pair((const void *)kfn, lds);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kfn, nt, lds) != hipSuccess || nb < 1) nb = 1;
    cache.emplace(key, nb);
    return nb;
}

// Dynamic LDS above 64 KiB needs the kernel's limit raised once (to the full
// 160 KiB, so later launches with any size are covered).
template <typename K> hipError_t raise_lds(K kfn, size_t bytes) {
    if (bytes <= 64 * 1024) return hipSuccess;
    static std::mutex mu;
    static std::unordered_set<const void *> done;
    std::lock_guard<std::mutex> g(mu);
    if (done.count((const void *)kfn)) return hipSuccess;
    hipError_t e = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess) done.insert((const void *)kfn);
    return e;
}

}  // namespace

// 0: automatic (fibre groups for n1 >= 16, LDS chunks below); 1: always the
// tiled GEMM; 2: flat fibre groups for n1 < 16 as well (1 and 2 for tests
// and timing; the LDS chunks measured 2% faster on a (16, 32, 16) chain)
int g_mode_kernel = 0;

int64_t zgemv_splits(int64_t m, int64_t n) {
    const int64_t rt = (m + 63) / 64;
    int64_t s = 1;
    while (rt * s < 1024 && n / (s * 2) >= 64) s *= 2;
    return s;
}

hipError_t launch_zgemm(qob_dtype dt, int64_t M, int64_t N, int64_t K, const void *A, int64_t sam, int64_t sak,
                        int64_t sa, const void *B, int64_t sbk, int64_t sbn, int64_t sb, void *C, int64_t scm,
                        int64_t scn, int64_t sc, int64_t batch, double2 alpha, double2 beta, hipStream_t s) {
    if (M <= 0 || N <= 0 || batch <= 0) return hipSuccess;
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    if (K <= 0) {
        // no product: C = beta C, element by element (beta = 0 writes zeros)
        return hipErrorInvalidValue;
    }
    const int64_t tmn = (M + GBM - 1) / GBM, tnn = (N + GBN - 1) / GBN;
    const int64_t nblk = tmn * tnn * batch;
    if (nblk > 0x7fffffff) return hipErrorInvalidValue;
    const size_t esz = dt == QOB_C64 ? 8 : 4;
    const size_t lds = (size_t)2 * 2 * GBK * ((GBM + GPAD) + (GBN + GPAD)) * esz;
    if (dt == QOB_C64) {
        auto k = k_zgemm<double>;
        hipError_t e = raise_lds(k, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(GNT), lds, s, M, N, K, (const double2 *)A, sam, sak, sa,
                           (const double2 *)B, sbk, sbn, sb, (double2 *)C, scm, scn, sc, tmn, tnn, alpha, beta, bz);
    } else {
        auto k = k_zgemm<float>;
        hipError_t e = raise_lds(k, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(GNT), lds, s, M, N, K, (const float2 *)A, sam, sak, sa,
                           (const float2 *)B, sbk, sbn, sb, (float2 *)C, scm, scn, sc, tmn, tnn, alpha, beta, bz);
    }
    return hipGetLastError();
}

hipError_t launch_zgemv(qob_dtype dt, const void *F, int64_t m, int64_t n, const void *x, void *y, void *part,
                        double2 alpha, double2 beta, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    const int64_t ns = zgemv_splits(m, n);
    const int64_t len = (n + ns - 1) / ns;
    if (ns > 1 && !part) return hipErrorInvalidValue;
    dim3 g((unsigned)((m + 63) / 64), (unsigned)ns);
    if (dt == QOB_C64) {
        hipLaunchKernelGGL(k_zgemv<double>, g, dim3(GV_NT), 0, s, (const double2 *)F, m, n, (const double2 *)x, len,
                           (double2 *)part, (double2 *)y, alpha, beta, bz);
        if (ns > 1)
            hipLaunchKernelGGL(k_zgemv_sum<double>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s,
                               (const double2 *)part, (int)ns, m, (double2 *)y, alpha, beta, bz);
    } else {
        hipLaunchKernelGGL(k_zgemv<float>, g, dim3(GV_NT), 0, s, (const float2 *)F, m, n, (const float2 *)x, len,
                           (double2 *)part, (float2 *)y, alpha, beta, bz);
        if (ns > 1)
            hipLaunchKernelGGL(k_zgemv_sum<float>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s,
                               (const double2 *)part, (int)ns, m, (float2 *)y, alpha, beta, bz);
    }
    return hipGetLastError();
}

hipError_t launch_zgemv_t(qob_dtype dt, const void *A, int64_t k, int64_t m, const void *x, void *y, double2 alpha,
                          double2 beta, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    const unsigned g = (unsigned)((m + 3) / 4);
    if (dt == QOB_C64)
        hipLaunchKernelGGL(k_zgemv_t<double>, dim3(g), dim3(256), 0, s, (const double2 *)A, k, m, (const double2 *)x,
                           (double2 *)y, alpha, beta, bz);
    else
        hipLaunchKernelGGL(k_zgemv_t<float>, dim3(g), dim3(256), 0, s, (const float2 *)A, k, m, (const float2 *)x,
                           (float2 *)y, alpha, beta, bz);
    return hipGetLastError();
}

// chunk of i3 blocks for k_mode_chunk: up to 32 KiB of planes (in + out),
// at least one block, and where that fits a whole number of fibre groups
// per wave (4 waves), else of single fibre groups
static int mode_chunk_blocks(int64_t n1, int64_t d_in, int64_t d_out, size_t esz) {
    const int64_t per = n1 * (d_in + d_out) * 2 * (int64_t)esz;  // plane bytes per i3 block
    int64_t c = std::max<int64_t>(1, (32 * 1024) / std::max<int64_t>(per, 1));
    const int64_t uw = 64 / std::gcd<int64_t>(n1, 64), u1 = 16 / std::gcd<int64_t>(n1, 16);
    if (c >= uw) c -= c % uw;
    else if (c >= u1) c -= c % u1;
    return (int)c;
}

size_t mode_workspace_bytes(int64_t d_out, int64_t d_in, int64_t n1, int64_t n3) {
    if (n1 == 1 && n3 == 1) {
        const int64_t ns = zgemv_splits(d_out, d_in);
        return ns > 1 ? (size_t)(ns * d_out) * 16 : 0;
    }
    return 0;
}

hipError_t launch_mode(qob_dtype dt, const void *F, int64_t d_out, int64_t d_in, const void *in, void *out,
                       int64_t n1, int64_t n3, double2 alpha, double2 beta, hipStream_t s, void *ws) {
    if (n1 <= 0 || n3 <= 0 || d_out <= 0) return hipSuccess;
    if (d_in <= 0) return hipErrorInvalidValue;
    const bool small = d_in <= 64 && d_out <= 64 && g_mode_kernel != 1;
    if (n1 == 1 && n3 == 1 && g_mode_kernel != 1) return launch_zgemv(dt, F, d_out, d_in, in, out, ws, alpha, beta, s);
    if (!small) {
        if (n1 == 1)  // out (d_out x n3) = F in
            return launch_zgemm(dt, d_out, n3, d_in, F, 1, d_out, 0, in, 1, d_in, 0, out, 1, d_out, 0, 1, alpha,
                                beta, s);
        if (n1 >= 64 || n3 == 1)  // per i3: out_i3 (n1 x d_out) = in_i3 (n1 x d_in) F^T
            return launch_zgemm(dt, n1, d_out, d_in, in, 1, n1, n1 * d_in, F, d_out, 1, 0, out, 1, n1, n1 * d_out,
                                n3, alpha, beta, s);
        // per i1: out[i1, :, :] (d_out x n3, strided) = F in[i1, :, :]
        return launch_zgemm(dt, d_out, n3, d_in, F, 1, d_out, 0, in, n1, n1 * d_in, 1, out, n1, n1 * d_out, 1, n1,
                            alpha, beta, s);
    }
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    const int nt16 = (int)((std::max(d_in, d_out) + 15) / 16);
    const size_t esz = dt == QOB_C64 ? 8 : 4;
    const size_t fbytes = (size_t)(4 * 4 * nt16) * (16 * nt16 + 1) * 2 * esz;
    const int ncu = device_cu_count();
    if (n1 >= 16 || g_mode_kernel == 2) {
        // n1 >= 16: groups of 16 consecutive i1; else of 16 consecutive fibres
        const int flat = n1 < 16;
        const int64_t nf = n1 * n3;
        const int64_t gpi = (n1 + 15) / 16, ng = flat ? (nf + 15) / 16 : gpi * n3;
#define QOB_MF(RR, NT)                                                                                          \
    do {                                                                                                        \
        auto k = k_mode_fib<RR, NT>;                                                                            \
        hipError_t e = raise_lds(k, fbytes);                                                                    \
        if (e != hipSuccess) return e;                                                                          \
        const int64_t grid = std::max<int64_t>(                                                                 \
            1, std::min<int64_t>((ng + 7) / 8, (int64_t)ncu * resident_per_cu(k, MODE_NT, fbytes)));           \
        hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(MODE_NT), fbytes, s, (const cplx<RR> *)F, (int)d_out, \
                           (int)d_in, (const cplx<RR> *)in, (cplx<RR> *)out, (int)n1, (int)gpi, ng, nf, alpha, beta, \
                           bz, flat);                                                                          \
        return hipGetLastError();                                                                               \
    } while (0)
        if (n1 > 0x7fffffff) return hipErrorInvalidValue;
        if (dt == QOB_C64) {
            if (nt16 == 1) QOB_MF(double, 1);
            if (nt16 == 2) QOB_MF(double, 2);
            if (nt16 == 3) QOB_MF(double, 3);
            QOB_MF(double, 4);
        } else {
            if (nt16 == 1) QOB_MF(float, 1);
            if (nt16 == 2) QOB_MF(float, 2);
            if (nt16 == 3) QOB_MF(float, 3);
            QOB_MF(float, 4);
        }
#undef QOB_MF
    }
    const int cpc = mode_chunk_blocks(n1, d_in, d_out, esz);
    const int64_t nch = (n3 + cpc - 1) / cpc;
    const size_t pin = (size_t)(cpc * n1 * d_in + ((cpc * n1 * d_in) >> 4) + 1);
    const size_t pout = (size_t)(cpc * n1 * d_out + ((cpc * n1 * d_out) >> 4) + 1);
    const size_t lds = fbytes + 2 * (pin + pout) * esz;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
#define QOB_MC(RR, NT)                                                                                        \
    do {                                                                                                      \
        auto k = k_mode_chunk<RR, NT>;                                                                        \
        hipError_t e = raise_lds(k, lds);                                                                     \
        if (e != hipSuccess) return e;                                                                        \
        const int64_t grid =                                                                                  \
            std::max<int64_t>(1, std::min<int64_t>(nch, (int64_t)ncu * resident_per_cu(k, CHUNK_NT, lds)));   \
        hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(CHUNK_NT), lds, s, (const cplx<RR> *)F, (int)d_out,  \
                           (int)d_in, (const cplx<RR> *)in, (cplx<RR> *)out, (int)n1, n3, cpc, nch, alpha, beta, bz); \
        return hipGetLastError();                                                                             \
    } while (0)
    if (dt == QOB_C64) {
        if (nt16 == 1) QOB_MC(double, 1);
        if (nt16 == 2) QOB_MC(double, 2);
        if (nt16 == 3) QOB_MC(double, 3);
        QOB_MC(double, 4);
    } else {
        if (nt16 == 1) QOB_MC(float, 1);
        if (nt16 == 2) QOB_MC(float, 2);
        if (nt16 == 3) QOB_MC(float, 3);
        QOB_MC(float, 4);
    }
#undef QOB_MC
}

}  // namespace qob
