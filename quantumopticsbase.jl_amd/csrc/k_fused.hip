// Fused dense LazyTensor application on MFMA: all mode products of a
// column in one pass.
//
// Reference: _tp_sum_matmul! (src/operators_lazytensor.jl:443-488) applies
// the factors one mode at a time, each _tp_matmul! a full pass over the
// state through temporaries, with permutations for middle modes
// (:333-404).  Here one workgroup owns one column (a Ket, or one column of a
// density matrix for the left application) at a time:
//   1. the column is loaded once into LDS as separate real and imaginary
//      planes, padded by one element every 16 (conflict-free MFMA operand
//      reads for every mode stride);
//   2. for each factor A (d x d, d a multiple of 16) the mode product
//      Y = A x_k X is a GEMM  (d x d) * (d x D/d): 16 x 16 output tiles on
//      v_mfma_{f64,f32}_16x16x4, accumulators in registers; all waves
//      finish reading before any tile is written back into the planes.
//      ComplexF64: four real MFMAs per complex k-step, as BLAS forms a
//      complex product, so each of the real and imaginary parts is as
//      accurate as the reference's; with one fibre group per wave the four
//      products go to four accumulators (Ar Xr, Ai Xi, Ar Xi, Ai Xr: four
//      independent MFMA chains; Re = first - second, Im = third + fourth at
//      the end), else to two (re*re - im*im, re*im + im*re).  ComplexF32:
//      three real MFMAs where the kernel has room for a third accumulator
//      (Gauss's product: P1 = Ar Xr, P2 = Ai Xi, P3 = (Ar + Ai)(Xr + Xi);
//      Re = P1 - P2, Im = P3 - P1 - P2; its error bound is normwise, the
//      imaginary part loses accuracy where it is much smaller than the real
//      part), else four;
//   3. the column is written once: out = alpha * Y + beta * out.
// ComplexF64 uses FP64 MFMA; ComplexF32 uses FP32 MFMA with FP32
// accumulation.  Columns are distributed over persistent workgroups.
//
// MFMA lane layouts: qob_mfma.h.
#include "qob_mfma.h"

#include <atomic>

namespace qob {

namespace {

__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }

constexpr int NTH = 512;  // threads per workgroup (8 waves); 1024 for one ComplexF32 shape (see the launcher)

// Each wave owns groups of 16 fibres; for a group it computes every 16-row
// output tile (ntm = d/16 of them) before writing the group back, so groups
// never overlap and no barrier is needed between compute and write-back.
// MAXNTM: d/16 of the widest factor; EPT: column elements per lane; GPW:
// fibre groups in flight per wave (more MFMA chains in flight, more registers).
// TILE = false: vector c is the contiguous column in + c * D (left
// application, Kets).  TILE = true (right application on a column-major
// matrix with leading dimension T.ld): vector c is the tile of 16 rows
// (row block c % T.nrb) by P = D / 16 consecutive columns (column block
// c / T.nrb), element e at row 16 (c % T.nrb) + (e & 15) and column
// (e >> 4) + P (c / T.nrb); the factors of the tile's modes act with the
// row index innermost.
struct FusedTile {
    int64_t ld = 0;
    int64_t nrb = 1;
    int P = 0;
};

// a[s] for a wave-uniform s through constant indices only: each a[i] is a
// scalar kernel-argument load, where a[s] would become a vector load that
// retires behind every global load issued before it.  The empty asm keeps
// each a[i] opaque (in an SGPR), so the compiler cannot fold the selects back
// into a[s].
__device__ __forceinline__ int mode_sel(const int (&a)[FUSED_MAXMODES], int s) {
    int r = a[0];
    asm("" : "+s"(r));
#pragma unroll
    for (int i = 1; i < FUSED_MAXMODES; ++i) {
        int v = a[i];
        asm("" : "+s"(v));
        r = (s == i) ? v : r;
    }
    return r;
}

template <bool TILE>
__device__ __forceinline__ int64_t fused_addr(const FusedTile &T, int D, int64_t c, int e) {
    if (!TILE) return c * (int64_t)D + e;
    const int64_t rb = c % T.nrb, q = c / T.nrb;
    return rb * 16 + (e & 15) + T.ld * ((int64_t)(e >> 4) + (int64_t)T.P * q);
}

// One mode of one column: each wave takes groups of 16 fibres (GPW at a
// time), forms every 16-row output tile of the group on MFMA and writes it
// back into the planes, or (last mode) to the output column.  DM > 0 fixes
// the mode size d at compile time; DM = 0 reads it at run time.  CM: how a
// complex k-step is formed (see the file header): CM_TWO four real products
// into two accumulators, CM_GAUSS three products, CM_FOUR four products
// into four accumulators.
enum { CM_TWO = 0, CM_GAUSS = 1, CM_FOUR = 2 };
template <typename R, int MAXNTM, int GPW, bool TILE, int DM, int CM, int NT>
__device__ __forceinline__ void fused_mode(R *re, R *im, const cplx<R> *Am, int ldam, int n1, int d_rt, int D,
                                           bool last, cplx<R> *__restrict__ out, int64_t col, const FusedTile &TL,
                                           cplx<R> alpha, cplx<R> beta, int beta_zero) {
    using T = cplx<R>;
    using V = typename mf<R>::v4;
    constexpr int NTM = DM > 0 ? DM / 16 : MAXNTM;  // accumulator tiles per group
    constexpr int KS = DM > 0 ? DM / 4 : MAXNTM * 4;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int d = DM > 0 ? DM : d_rt;
    const int ntm = d >> 4, ngroups = (D / d) >> 4;
    const int nks = d >> 2;
    for (int g0 = wave; g0 < ngroups; g0 += (NT / 64) * GPW) {
        // CM_GAUSS: accs = P3; CM_FOUR: accr = Ar Xr, acci = Ar Xi,
        // accs = Ai Xi, acct = Ai Xr
        V accr[GPW][NTM], acci[GPW][NTM], accs[GPW][NTM], acct[GPW][NTM];
        int base[GPW];
#pragma unroll
        for (int q = 0; q < GPW; ++q) {
            const int g = g0 + q * (NT / 64);
            const int f = g * 16 + lr;
            const int i1 = f % n1, i3 = f / n1;
            base[q] = i1 + n1 * d * i3;
#pragma unroll
            for (int mt = 0; mt < NTM; ++mt) {
                accr[q][mt] = V{0, 0, 0, 0};
                acci[q][mt] = V{0, 0, 0, 0};
                accs[q][mt] = V{0, 0, 0, 0};
                acct[q][mt] = V{0, 0, 0, 0};
            }
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            if (ks >= nks) break;
            R xr[GPW], xi[GPW];
#pragma unroll
            for (int q = 0; q < GPW; ++q) {
                const int idx = (g0 + q * (NT / 64) < ngroups) ? pad(base[q] + n1 * (4 * ks + lk)) : 0;
                xr[q] = re[idx];
                xi[q] = im[idx];
            }
#pragma unroll
            for (int mt = 0; mt < NTM; ++mt) {
                if (mt < ntm) {
                    const T a = Am[(mt * 16 + lr) + (4 * ks + lk) * ldam];
#pragma unroll
                    for (int q = 0; q < GPW; ++q) {
                        if constexpr (CM == CM_GAUSS) {
                            accr[q][mt] = mf<R>::mma(a.x, xr[q], accr[q][mt]);
                            acci[q][mt] = mf<R>::mma(a.y, xi[q], acci[q][mt]);
                            accs[q][mt] = mf<R>::mma(a.x + a.y, xr[q] + xi[q], accs[q][mt]);
                        } else if constexpr (CM == CM_FOUR) {
                            accr[q][mt] = mf<R>::mma(a.x, xr[q], accr[q][mt]);
                            acci[q][mt] = mf<R>::mma(a.x, xi[q], acci[q][mt]);
                            accs[q][mt] = mf<R>::mma(a.y, xi[q], accs[q][mt]);
                            acct[q][mt] = mf<R>::mma(a.y, xr[q], acct[q][mt]);
                        } else {
                            accr[q][mt] = mf<R>::mma(a.x, xr[q], accr[q][mt]);
                            accr[q][mt] = mf<R>::mma(a.y, -xi[q], accr[q][mt]);
                            acci[q][mt] = mf<R>::mma(a.x, xi[q], acci[q][mt]);
                            acci[q][mt] = mf<R>::mma(a.y, xr[q], acci[q][mt]);
                        }
                    }
                }
            }
        }
        if constexpr (CM == CM_GAUSS) {
            // (P1, P2, P3) -> (Re, Im)
#pragma unroll
            for (int q = 0; q < GPW; ++q) {
#pragma unroll
                for (int mt = 0; mt < NTM; ++mt) {
                    const V p1 = accr[q][mt], p2 = acci[q][mt];
                    accr[q][mt] = p1 - p2;
                    acci[q][mt] = accs[q][mt] - p1 - p2;
                }
            }
        } else if constexpr (CM == CM_FOUR) {
#pragma unroll
            for (int q = 0; q < GPW; ++q) {
#pragma unroll
                for (int mt = 0; mt < NTM; ++mt) {
                    accr[q][mt] = accr[q][mt] - accs[q][mt];
                    acci[q][mt] = acci[q][mt] + acct[q][mt];
                }
            }
        }
        if (!last) {
            // the group's inputs are no longer needed by anyone: write back
#pragma unroll
            for (int q = 0; q < GPW; ++q) {
                if (g0 + q * (NT / 64) < ngroups) {
#pragma unroll
                    for (int mt = 0; mt < NTM; ++mt) {
                        if (mt < ntm) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int a = mt * 16 + mf<R>::row(lane, r);
                                const int idx = pad(base[q] + n1 * a);
                                re[idx] = accr[q][mt][r];
                                im[idx] = acci[q][mt][r];
                            }
                        }
                    }
                }
            }
        } else {
            // last mode: the tiles go from the accumulators straight to the
            // output column (out = alpha y + beta out); the stores drain
            // while the next column is being computed
#pragma unroll
            for (int q = 0; q < GPW; ++q) {
                if (g0 + q * (NT / 64) < ngroups) {
#pragma unroll
                    for (int mt = 0; mt < NTM; ++mt) {
                        if (mt < ntm) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int e = base[q] + n1 * (mt * 16 + mf<R>::row(lane, r));
                                T v = cmul(alpha, mk((R)accr[q][mt][r], (R)acci[q][mt][r]));
                                const int64_t ge = fused_addr<TILE>(TL, D, col, e);
                                if (!beta_zero) v = cfma(beta, out[ge], v);
                                st_stream(out + ge, v);
                            }
                        }
                    }
                }
            }
        }
    }
}

// RES = true: every factor stays in LDS for the whole launch (mode s at
// offset sum_{t<s} d_t^2, ld = d), staged once per workgroup.  The mode loop
// then issues no global loads besides the next column's prefetch, so that
// prefetch stays in flight across all modes (loads retire in order: a
// factor load issued after it would have to wait for the whole column), and
// each mode needs one barrier instead of two.
template <typename R, int MAXNTM, int EPT, int GPW, bool TILE, bool RES = false, int NT = NTH>
__global__ void __launch_bounds__(NT) k_lt_fused(FusedModes M, const cplx<R> *__restrict__ in,
                                                   cplx<R> *__restrict__ out, int64_t ncols, FusedTile TL,
                                                   double2 alpha_, double2 beta_, int beta_zero) {
    using T = cplx<R>;
    extern __shared__ unsigned char lds_raw[];
    R *re = (R *)lds_raw;
    const int D = M.D;
    const int Dp = D + (D >> 4);
    R *im = re + Dp;
    T *Al = (T *)(im + Dp + (Dp & 1));  // factor of the current mode, ld = d + 1 (RES: all factors)
    const int tid = threadIdx.x;
    const T alpha = mk((R)alpha_.x, (R)alpha_.y), beta = mk((R)beta_.x, (R)beta_.y);

    // factor of the next mode, loaded into registers one mode ahead so the
    // load latency hides behind the current mode's MFMAs
    constexpr int APRE = (16 * MAXNTM) * (16 * MAXNTM) / NT > 0 ? (16 * MAXNTM) * (16 * MAXNTM) / NT : 1;
    T apre[APRE];
    if constexpr (RES) {
        // ordered before every read of the factors by the first mode's barrier
        int off = 0;
        for (int s = 0; s < M.nmodes; ++s) {
            const T *__restrict__ As = (const T *)M.A[s];
            const int dd = M.d[s] * M.d[s];
            for (int e = tid; e < dd; e += NT) Al[off + e] = As[e];
            off += dd;
        }
    } else {
        const T *__restrict__ A0 = (const T *)M.A[0];
        const int dd = M.d[0] * M.d[0];
#pragma unroll
        for (int k = 0; k < APRE; ++k) {
            const int e = tid + k * NT;
            apre[k] = (e < dd) ? A0[e] : czero<T>();
        }
    }

    T pre[EPT];
    int64_t col = blockIdx.x;
    if (col < ncols) {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = tid + k * NT;
            if constexpr (RES) pre[k] = in[fused_addr<TILE>(TL, D, col, e < D ? e : D - 1)];
            else if (e < D) pre[k] = in[fused_addr<TILE>(TL, D, col, e)];
        }
    }
    for (; col < ncols; col += gridDim.x) {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = tid + k * NT;
            if constexpr (RES) {
                // branch-free, so the wait for pre[k] counts exactly and does
                // not also wait for the previous column's output stores; lanes
                // past D write slot 16, a padding hole no element maps to
                const int idx = e < D ? pad(e) : 16;
                re[idx] = pre[k].x;
                im[idx] = pre[k].y;
            } else if (e < D) {
                re[pad(e)] = pre[k].x;
                im[pad(e)] = pre[k].y;
            }
        }
        // issue the next column's loads now; they complete during the MFMAs
        const int64_t nxt = col + gridDim.x;
        if (nxt < ncols) {
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const int e = tid + k * NT;
                if constexpr (RES) {
                    // unconditional (clamped) loads: no branch between them and
                    // their use, so the wait before the use counts exactly
                    pre[k] = in[fused_addr<TILE>(TL, D, nxt, e < D ? e : D - 1)];
                } else if (e < D) {
                    pre[k] = in[fused_addr<TILE>(TL, D, nxt, e)];
                }
            }
        }
        int aoff = 0;  // RES: offset of the current mode's factor

        // FP32: Gauss's product where its third accumulator adds no register
        // spills (factors up to 32 wide, up to 16 elements per lane, up to
        // two fibre groups per wave).  FP64: four products, into four
        // accumulators with resident factors up to 32 wide and one fibre
        // group per wave
        constexpr int CM = (sizeof(R) == 4 && MAXNTM <= 2 && EPT <= 16 && GPW <= 2) ? CM_GAUSS
                           : (sizeof(R) == 8 && MAXNTM <= 2 && GPW == 1 && RES)   ? CM_FOUR
                                                                                  : CM_TWO;
        for (int s = 0; s < M.nmodes; ++s) {
            const int n1 = RES ? mode_sel(M.n1, s) : M.n1[s], d = RES ? mode_sel(M.d, s) : M.d[s];
            const int lda = d + 1;
            __syncthreads();  // previous mode written back, column loaded
            if constexpr (!RES) {
#pragma unroll
            for (int k = 0; k < APRE; ++k) {
                const int e = tid + k * NT;
                if (e < d * d) {
                    const int i = e % d, kk = e / d;
                    Al[i + kk * lda] = apre[k];
                }
            }
            {
                // next factor: this column's next mode, else the first mode
                // of the next column
                const int sn = (s + 1 < M.nmodes) ? s + 1 : 0;
                const T *__restrict__ An = (const T *)M.A[sn];
                const int dd = M.d[sn] * M.d[sn];
#pragma unroll
                for (int k = 0; k < APRE; ++k) {
                    const int e = tid + k * NT;
                    apre[k] = (e < dd) ? An[e] : czero<T>();
                }
            }
            __syncthreads();
            }
            const T *Am = RES ? Al + aoff : Al;
            const int ldam = RES ? d : lda;
            aoff += d * d;
            const bool last = s + 1 == M.nmodes;
            // mode size as a template argument: the k-step and tile loops
            // unroll to a fixed MFMA sequence with no branches in between
            // (factors up to 32 wide).  Kept at run-time sizes: kernels with
            // 64-wide factors and the ComplexF64 tile variant, which spill
            // further with fixed sizes (the tile variant measured 1.30
            // against 1.18 ms).
            if constexpr (RES && MAXNTM <= 2 && !(TILE && sizeof(R) == 8 && GPW == 2)) {
                if (d == 16) {
                    fused_mode<R, MAXNTM, GPW, TILE, 16, CM, NT>(re, im, Am, ldam, n1, d, D, last, out, col, TL, alpha, beta, beta_zero);
                } else {
                    fused_mode<R, MAXNTM, GPW, TILE, 32, CM, NT>(re, im, Am, ldam, n1, d, D, last, out, col, TL, alpha, beta, beta_zero);
                }
            } else {
                fused_mode<R, MAXNTM, GPW, TILE, 0, CM, NT>(re, im, Am, ldam, n1, d, D, last, out, col, TL, alpha, beta, beta_zero);
            }
        }
        __syncthreads();  // every wave is done reading this column's planes
    }
}

}  // namespace

// fibre groups in flight per wave: 1, 2 or 4; 0 (automatic): one for
// ComplexF64 with resident factors up to 32 wide (four accumulators), else two
int g_fused_gpw = 0;
int g_fused_tiles = 1;     // 1: right applications run the fused kernel on 16-row tiles
int g_fused_resident = 1;  // 1: factors resident in LDS when planes + factors fit (two fibre groups per wave)

// LDS bytes for a column of D elements (padded planes)
static size_t fused_lds(qob_dtype dt, int D, int dmax) {
    const size_t r = (dt == QOB_C64 ? 8 : 4);
    const size_t dp = (size_t)(D + (D >> 4));
    return (2 * dp + (dp & 1)) * r + (size_t)dmax * (dmax + 1) * 2 * r;
}

// resident factors: planes plus every factor at ld = d
static size_t fused_lds_res(qob_dtype dt, const FusedModes &M) {
    const size_t r = (dt == QOB_C64 ? 8 : 4);
    const size_t dp = (size_t)(M.D + (M.D >> 4));
    size_t a = 0;
    for (int s = 0; s < M.nmodes; ++s) a += (size_t)M.d[s] * M.d[s];
    return (2 * dp + (dp & 1)) * r + a * 2 * r;
}

static int fused_dmax(const FusedModes &M) {
    int m = 0;
    for (int i = 0; i < M.nmodes; ++i) m = std::max(m, M.d[i]);
    return m;
}

bool fused_supported(qob_dtype dt, const FusedModes &M) {
    if (M.nmodes < 1 || M.nmodes > FUSED_MAXMODES || M.D <= 0 || (M.D & 255)) return false;
    const int dmax = (dt == QOB_C64) ? 8192 : 16384;
    if (M.D > dmax) return false;
    if (fused_lds(dt, M.D, fused_dmax(M)) > 160 * 1024) return false;
    int64_t prodd = 1;
    for (int s = 0; s < M.nmodes; ++s) {
        const int d = M.d[s];
        if (d != 16 && d != 32 && d != 64) return false;  // d/16 divides the 8 waves
        if (((M.D / d) & 15) != 0) return false;
        if ((int64_t)M.n1[s] * d > M.D) return false;
        prodd *= d;
    }
    return prodd <= M.D;
}

hipError_t launch_lt_fused(qob_dtype dt, const FusedModes &M, const void *in, void *out, int64_t ncols,
                           double2 alpha, double2 beta, hipStream_t s, int64_t tile_ld, int64_t tile_nrb) {
    // tile_ld > 0: right application in tiles of 16 rows (see FusedTile)
    FusedTile TL;
    if (tile_ld > 0) {
        if (M.D % 16 || tile_nrb < 1 || tile_ld < 16 * tile_nrb) return hipErrorInvalidValue;
        TL.ld = tile_ld;
        TL.nrb = tile_nrb;
        TL.P = M.D / 16;
    }
    if (ncols == 0) return hipSuccess;
    if (!fused_supported(dt, M)) return hipErrorInvalidValue;
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    const int maxd = fused_dmax(M);
    const bool fits_res = g_fused_resident && fused_lds_res(dt, M) <= 160 * 1024;
    const int gpw = g_fused_gpw ? g_fused_gpw : (dt == QOB_C64 && maxd <= 32 && fits_res ? 1 : 2);
    const bool res = fits_res && gpw <= 2;
    // automatic shape, ComplexF32 left application of resident factors up to
    // 32 wide on columns of 4096..8192: 1024 threads with one fibre group per
    // wave (16 waves of one workgroup share each column's phases); longer
    // columns would spill at 1024 threads
    const bool wide = g_fused_gpw == 0 && dt == QOB_C32 && res && tile_ld == 0 && maxd <= 32 && M.D >= 4096 &&
                      M.D <= 8192;
    const size_t lds = res ? fused_lds_res(dt, M) : fused_lds(dt, M.D, maxd);
    const int ept = (M.D + NTH - 1) / NTH;
    const int ncu = device_cu_count();
    const int64_t per_cu = std::max<int64_t>(1, (160 * 1024) / (int64_t)lds);
    const unsigned g = (unsigned)std::min<int64_t>(ncols, (int64_t)ncu * per_cu);
#define QOB_FU3(RR, NTM, EP, GP, TI)                                                                             \
    do {                                                                                                     \
        auto kfn = k_lt_fused<RR, NTM, EP, GP, TI>;                                                          \
        static std::atomic<int> lds_set{0}; /* raise the limit once per kernel */                            \
        if (!lds_set.load()) {                                                                               \
            hipError_t e_ = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                160 * 1024);                                                 \
            if (e_ != hipSuccess) return e_;                                                                 \
            lds_set.store(1);                                                                                \
        }                                                                                                    \
        hipLaunchKernelGGL(kfn, dim3(g), dim3(NTH), lds, s, M, (const cplx<RR> *)in, (cplx<RR> *)out, ncols, \
                           TL, alpha, beta, bz);                                                             \
        return hipGetLastError();                                                                            \
    } while (0)
#define QOB_FU3R(RR, NTM, EP, GP, TI)                                                                             \
    do {                                                                                                     \
        auto kfn = k_lt_fused<RR, NTM, EP, GP, TI, true>;                                                          \
        static std::atomic<int> lds_set{0}; /* raise the limit once per kernel */                            \
        if (!lds_set.load()) {                                                                               \
            hipError_t e_ = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                160 * 1024);                                                 \
            if (e_ != hipSuccess) return e_;                                                                 \
            lds_set.store(1);                                                                                \
        }                                                                                                    \
        hipLaunchKernelGGL(kfn, dim3(g), dim3(NTH), lds, s, M, (const cplx<RR> *)in, (cplx<RR> *)out, ncols, \
                           TL, alpha, beta, bz);                                                             \
        return hipGetLastError();                                                                            \
    } while (0)
#define QOB_FUW(NTM, EP)                                                                                       \
    do {                                                                                                     \
        auto kfn = k_lt_fused<float, NTM, EP, 1, false, true, 1024>;                                         \
        static std::atomic<int> lds_set{0}; /* raise the limit once per kernel */                            \
        if (!lds_set.load()) {                                                                               \
            hipError_t e_ = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                160 * 1024);                                                 \
            if (e_ != hipSuccess) return e_;                                                                 \
            lds_set.store(1);                                                                                \
        }                                                                                                    \
        /* persistent grid of the workgroups that are resident at once */                                \
        static std::atomic<int> nres{0};                                                                     \
        if (!nres.load()) {                                                                                  \
            int nb_ = 0;                                                                                     \
            hipError_t e_ = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_, kfn, 1024, lds);              \
            if (e_ != hipSuccess) return e_;                                                                 \
            nres.store(nb_ > 0 ? nb_ : 1);                                                                   \
        }                                                                                                    \
        const unsigned gw = (unsigned)std::min<int64_t>(ncols, (int64_t)ncu * nres.load());                 \
        hipLaunchKernelGGL(kfn, dim3(gw), dim3(1024), lds, s, M, (const cplx<float> *)in, (cplx<float> *)out, \
                           ncols, TL, alpha, beta, bz);                                                      \
        return hipGetLastError();                                                                            \
    } while (0)
#define QOB_FU(RR, NTM, EP)                                          \
    do {                                                             \
        if (res) {                                                   \
            if (gpw == 1) {                                          \
                if (tile_ld > 0) QOB_FU3R(RR, NTM, EP, 1, true);     \
                QOB_FU3R(RR, NTM, EP, 1, false);                     \
            }                                                        \
            if (tile_ld > 0) QOB_FU3R(RR, NTM, EP, 2, true);         \
            QOB_FU3R(RR, NTM, EP, 2, false);                         \
        }                                                            \
        if (tile_ld > 0) QOB_FU3(RR, NTM, EP, 2, true);              \
        else if (gpw == 1) QOB_FU3(RR, NTM, EP, 1, false);           \
        else if (gpw == 4) QOB_FU3(RR, NTM, EP, 4, false);           \
        else QOB_FU3(RR, NTM, EP, 2, false);                         \
    } while (0)
    const int ntm = maxd / 16;
    if (dt == QOB_C64) {
        if (ept <= 4) { if (ntm <= 2) QOB_FU(double, 2, 4); else QOB_FU(double, 4, 4); }
        if (ept <= 8) { if (ntm <= 2) QOB_FU(double, 2, 8); else QOB_FU(double, 4, 8); }
        if (ept <= 16) { if (ntm <= 2) QOB_FU(double, 2, 16); else QOB_FU(double, 4, 16); }
    } else {
        if (wide) QOB_FUW(2, 8);
        if (ept <= 8) { if (ntm <= 2) QOB_FU(float, 2, 8); else QOB_FU(float, 4, 8); }
        if (ept <= 16) { if (ntm <= 2) QOB_FU(float, 2, 16); else QOB_FU(float, 4, 16); }
        if (ept <= 32) { if (ntm <= 2) QOB_FU(float, 2, 32); else QOB_FU(float, 4, 32); }
    }
    return hipErrorInvalidValue;
#undef QOB_FU
#undef QOB_FUW
#undef QOB_FU3R
#undef QOB_FU3
    return hipGetLastError();
}

}  // namespace qob
