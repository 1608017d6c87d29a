// Dense LazyTensor with factors on sites 1, 2 (and 3) of a column: modes 1
// and 2 chained in registers, one LDS exchange for mode 3.
//
// Reference: _tp_sum_matmul! (src/operators_lazytensor.jl:443-488) applies the
// factors one mode at a time, each a full pass over the state.  For a column
// x viewed as [i1 (16), i2 (d2), i3 (R)] (site 1 fastest) with factors A1
// (16 x 16) on site 1, A2 (d2 x d2) on site 2 and optionally A3 (R x R) on
// site 3, this kernel forms y = (A3 (x) A2 (x) A1) x as follows.  A block is
// the 16 x d2 slice M = x[:, :, i3] (contiguous).
//   mode 1, transposed: D1 = M^T A1^T (rows i2, columns i1'), MFMA A-operand
//     the block read straight from HBM, B-operand A1^T from registers;
//   mode 2: D2 = A2 D1, whose B-operand for k-step k is exactly register k % 4
//     of D1 tile k / 4 (the f64 D tile row (lane/16) + 4 r, and the f32 row
//     4 (lane/16) + r, are matched by the A2 fragments' k index), so the
//     block never leaves registers between the two modes;
//   mode 3 mixes the blocks: every wave writes its blocks' results into LDS
//     planes once, and after a barrier the mode-3 products read them back and
//     store the column to HBM.  Without a site-3 factor the blocks go from
//     registers straight to HBM.
// Per column that is one HBM read, one HBM write, one LDS write and one LDS
// read of the column, and two barriers (the fused column kernel of
// k_fused.hip writes the column back to LDS after every mode, with four
// barriers).  Every complex product is four real MFMAs (qob_mfma.h).
#include "qob_mfma.h"

#include <atomic>

namespace qob {

namespace {

__device__ __forceinline__ int pad16r(int i) { return i + (i >> 4); }

constexpr int REG_NT = 512;  // 8 waves

// k-index of mode-2 k-step ks, for the lane's k-row lk: the row of D1 tile
// ks / 4 that register ks % 4 holds
template <typename R> __device__ __forceinline__ int d1_row(int ks, int lk) {
    const int t = ks >> 2, r = ks & 3;
    return 16 * t + mf<R>::row(lk * 16, r);
}

// The mode-1 A-fragments of one block (16 x D2 elements at xb): element
// (i1 = 4 ks + lane / 16, i2 = 16 t + lane % 16), read straight from HBM.
template <typename R, int D2>
__device__ __forceinline__ void block_load(const cplx<R> *__restrict__ xb, int lane, cplx<R> (&x)[D2 / 16][4]) {
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int t = 0; t < D2 / 16; ++t)
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) x[t][ks] = xb[(4 * ks + lk) + 16 * (16 * t + lr)];
}

// Modes 1 and 2 of one block from its loaded fragments: on return d2r/d2i
// hold tile u of D2 = A2 (M^T A1^T), element (i1' = lane % 16,
// i2' = 16 u + mf<R>::row(lane, r)) in register r.
template <typename R, int D2, int LDA2>
__device__ __forceinline__ void block_modes12(const cplx<R> (&x)[D2 / 16][4], const cplx<R> (&a1f)[4],
                                              const cplx<R> *A2s, int lane, typename mf<R>::v4 (&d2r)[D2 / 16],
                                              typename mf<R>::v4 (&d2i)[D2 / 16]) {
    using T = cplx<R>;
    using V = typename mf<R>::v4;
    constexpr int NT2 = D2 / 16, KS2 = D2 / 4;
    const int lr = lane & 15, lk = lane >> 4;
    // mode 1: D1[t] = M^T A1^T over rows i2 = 16 t + ...
    V d1r[NT2], d1i[NT2];
#pragma unroll
    for (int t = 0; t < NT2; ++t) {
        d1r[t] = V{0, 0, 0, 0};
        d1i[t] = V{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) cmma4<R>(x[t][ks], a1f[ks], d1r[t], d1i[t]);
    }
    // mode 2: D2[u] = A2 D1 (rows i2' = 16 u + ..., columns i1')
#pragma unroll
    for (int u = 0; u < NT2; ++u) {
        d2r[u] = V{0, 0, 0, 0};
        d2i[u] = V{0, 0, 0, 0};
    }
#pragma unroll
    for (int ks = 0; ks < KS2; ++ks) {
        const int t = ks >> 2, r = ks & 3;
        const T bv = mk((R)d1r[t][r], (R)d1i[t][r]);
        const int i2 = d1_row<R>(ks, lk);
#pragma unroll
        for (int u = 0; u < NT2; ++u) cmma4<R>(A2s[(16 * u + lr) + i2 * LDA2], bv, d2r[u], d2i[u]);
    }
}

// D2 = 16 x d2 blocks per column, nblk blocks per column; D3 > 0: the site-3
// factor is nblk x nblk (nblk == D3), 0: no site-3 factor.
template <typename R, int D2, int D3>
__global__ void __launch_bounds__(REG_NT) k_lt_reg(const cplx<R> *__restrict__ A1, const cplx<R> *__restrict__ A2,
                                                    const cplx<R> *__restrict__ A3, const cplx<R> *__restrict__ in,
                                                    cplx<R> *__restrict__ out, int64_t ncols, int nblk,
                                                    double2 alpha_, double2 beta_, int beta_zero) {
    using T = cplx<R>;
    using V = typename mf<R>::v4;
    constexpr int BLK = 16 * D2;          // elements per block
    constexpr int NT2 = D2 / 16;          // D2 tiles per block
    constexpr int WPB = REG_NT / 64;
    constexpr int LDA2 = D2 + 1;
    extern __shared__ unsigned char lds_raw[];
    T *A2s = (T *)lds_raw;                                   // A2, ld = d2 + 1
    const int D = BLK * nblk;
    const int Dp = pad16r(D) + 1;
    R *yre = (R *)(A2s + D2 * LDA2);
    R *yim = yre + Dp;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const T alpha = mk((R)alpha_.x, (R)alpha_.y), beta = mk((R)beta_.x, (R)beta_.y);

    for (int e = tid; e < D2 * D2; e += REG_NT) A2s[(e % D2) + (e / D2) * LDA2] = A2[e];
    // A1^T B-fragments (k = i1, j = i1'): A1[i1' + 16 i1]
    T a1f[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) a1f[ks] = A1[lr + 16 * (4 * ks + lk)];
    // A3 A-fragments (i = k3', k = k3)
    constexpr int NT3 = D3 > 0 ? D3 / 16 : 1, KS3 = D3 > 0 ? D3 / 4 : 1;
    T a3f[NT3][KS3];
    if constexpr (D3 > 0) {
#pragma unroll
        for (int v = 0; v < NT3; ++v)
#pragma unroll
            for (int ks = 0; ks < KS3; ++ks) a3f[v][ks] = A3[(16 * v + lr) + (int64_t)D3 * (4 * ks + lk)];
    }
    __syncthreads();  // A2 staged

    // ComplexF64: global loads run one block ahead, a block's fragments
    // requested before the previous block's MFMAs (and the first block of the
    // next column before this column's mode 3).  ComplexF32 loads each block
    // just before its MFMAs: the registers of a second block would halve the
    // workgroups that fit on a CU (measured slower).
    constexpr bool PIPE = sizeof(R) == 8;
    T xa[NT2][4], xb[NT2][4];
    int64_t col = blockIdx.x;
    if (PIPE && col < ncols && wave < nblk)
        block_load<R, D2>(in + col * (int64_t)D + (int64_t)wave * BLK, lane, xa);
    for (; col < ncols; col += gridDim.x) {
        const T *xc = in + col * (int64_t)D;
        T *oc = out + col * (int64_t)D;
        const int64_t ncol = col + gridDim.x;
        V d2r[NT2], d2i[NT2];
        if constexpr (D3 == 0) {
            // no site-3 factor: every block goes from registers to HBM
            for (int blk = wave; blk < nblk; blk += WPB) {
                if (!PIPE) block_load<R, D2>(xc + (int64_t)blk * BLK, lane, xa);
                block_modes12<R, D2, LDA2>(xa, a1f, A2s, lane, d2r, d2i);
                if (PIPE && blk + WPB < nblk) block_load<R, D2>(xc + (int64_t)(blk + WPB) * BLK, lane, xa);
                else if (PIPE && ncol < ncols)
                    block_load<R, D2>(in + ncol * (int64_t)D + (int64_t)wave * BLK, lane, xa);
#pragma unroll
                for (int u = 0; u < NT2; ++u)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        T *o = oc + (int64_t)blk * BLK + lr + 16 * (16 * u + mf<R>::row(lane, r));
                        T v = cmul(alpha, mk((R)d2r[u][r], (R)d2i[u][r]));
                        if (!beta_zero) v = cfma(beta, *o, v);
                        st_stream(o, v);
                    }
            }
        } else {
            // the wave's first block is formed while other waves may still be
            // reading the previous column's planes; it is written after the
            // barrier, the wave's other blocks straight after they are formed
            if (wave < nblk) {
                if (!PIPE) block_load<R, D2>(xc + (int64_t)wave * BLK, lane, xa);
                block_modes12<R, D2, LDA2>(xa, a1f, A2s, lane, d2r, d2i);
                if (PIPE && wave + WPB < nblk) block_load<R, D2>(xc + (int64_t)(wave + WPB) * BLK, lane, xb);
            }
            __syncthreads();  // every wave is done reading the previous column's planes
            for (int blk = wave; blk < nblk; blk += WPB) {
                if (blk != wave) {
                    if (!PIPE) block_load<R, D2>(xc + (int64_t)blk * BLK, lane, xb);
                    block_modes12<R, D2, LDA2>(xb, a1f, A2s, lane, d2r, d2i);
                    if (PIPE && blk + WPB < nblk) block_load<R, D2>(xc + (int64_t)(blk + WPB) * BLK, lane, xb);
                }
#pragma unroll
                for (int u = 0; u < NT2; ++u)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int e = blk * BLK + lr + 16 * (16 * u + mf<R>::row(lane, r));
                        yre[pad16r(e)] = d2r[u][r];
                        yim[pad16r(e)] = d2i[u][r];
                    }
            }
            __syncthreads();  // the column's planes are complete
            if (PIPE && ncol < ncols && wave < nblk)
                block_load<R, D2>(in + ncol * (int64_t)D + (int64_t)wave * BLK, lane, xa);
            // mode 3: Z^T (k3' x e) = A3 Y^T over groups of 16 elements e
            if constexpr (PIPE && NT3 == 1) {
                // ComplexF64 with one output tile per group: every wave runs
                // the same compile-time number of groups, so the wait for the
                // next column's loads (issued above, before this mode's
                // stores) counts all of the stores that follow them.  With a
                // run-time trip count the compiler assumes one group's stores
                // and that wait also waited for 12 of the 16 (1% of the
                // kernel's time).  The scheduling fence keeps one group's
                // registers live at a time; with two tiles per group (a
                // 32-wide site 3) the unrolled form spills, so that shape
                // keeps the loop.
                auto group3 = [&](int g) {
                    const int e0 = 16 * g;
                    V zr[NT3], zi[NT3];
#pragma unroll
                    for (int v = 0; v < NT3; ++v) {
                        zr[v] = V{0, 0, 0, 0};
                        zi[v] = V{0, 0, 0, 0};
                    }
#pragma unroll
                    for (int ks = 0; ks < KS3; ++ks) {
                        const int idx = pad16r(e0 + lr + BLK * (4 * ks + lk));
                        const T yv = mk(yre[idx], yim[idx]);
#pragma unroll
                        for (int v = 0; v < NT3; ++v) cmma4<R>(a3f[v][ks], yv, zr[v], zi[v]);
                    }
#pragma unroll
                    for (int v = 0; v < NT3; ++v)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            T *o = oc + e0 + lr + (int64_t)BLK * (16 * v + mf<R>::row(lane, r));
                            T val = cmul(alpha, mk((R)zr[v][r], (R)zi[v][r]));
                            if (!beta_zero) val = cfma(beta, *o, val);
                            st_stream(o, val);
                        }
                };
                static_assert((BLK / 16) % WPB == 0, "mode-3 groups per wave");
#pragma unroll
                for (int q = 0; q < BLK / 16 / WPB; ++q) {
                    __builtin_amdgcn_sched_barrier(0);
                    group3(wave + q * WPB);
                }
            } else {
                // the body of group3, written out: calling the lambda here
                // costs the ComplexF32 forms 4-9 more VGPRs
                for (int g = wave; g < BLK / 16; g += WPB) {
                    const int e0 = 16 * g;
                    V zr[NT3], zi[NT3];
#pragma unroll
                    for (int v = 0; v < NT3; ++v) {
                        zr[v] = V{0, 0, 0, 0};
                        zi[v] = V{0, 0, 0, 0};
                    }
#pragma unroll
                    for (int ks = 0; ks < KS3; ++ks) {
                        const int idx = pad16r(e0 + lr + BLK * (4 * ks + lk));
                        const T yv = mk(yre[idx], yim[idx]);
#pragma unroll
                        for (int v = 0; v < NT3; ++v) cmma4<R>(a3f[v][ks], yv, zr[v], zi[v]);
                    }
#pragma unroll
                    for (int v = 0; v < NT3; ++v)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            T *o = oc + e0 + lr + (int64_t)BLK * (16 * v + mf<R>::row(lane, r));
                            T val = cmul(alpha, mk((R)zr[v][r], (R)zi[v][r]));
                            if (!beta_zero) val = cfma(beta, *o, val);
                            st_stream(o, val);
                        }
                }
            }
        }
    }
}

// Right application on tiles of 16 rows (R = rho * LT on a column-major
// rho of leading dimension ld): tile c holds rows 16 (c % nrb) + 0..15 of the
// D2 * D3 consecutive columns starting at (c / nrb) * D2 * D3, viewed as
// [r (16), i2 (D2), i3 (D3)] with element (r, i2, i3) at
// base + r + ld (i2 + D2 i3).  The factors act on i2 and i3 (the rows are a
// batch index), so this is the column kernel above without a site-1 factor:
// a block (16 rows x D2 columns, fixed i3) is loaded straight into mode 2's
// B-operand layout (k = i2 along lane / 16, the 16 rows along lane % 16:
// four 256-byte runs per load instruction in ComplexF64), mode 2 runs on it
// in registers and mode 3 goes through the LDS planes as above.  Mode 2 is
// short (one MFMA k-chain per block), so a wave's blocks of the next tile
// are all loaded while this tile's mode 3 runs; the site-3 factor sits in
// LDS beside the planes to leave registers for them.
template <typename R, int D2, int D3>
__global__ void __launch_bounds__(REG_NT) k_lt_reg_tile(const cplx<R> *__restrict__ A2, const cplx<R> *__restrict__ A3,
                                                         const cplx<R> *__restrict__ in, cplx<R> *__restrict__ out,
                                                         int64_t ntiles, int64_t nrb, int64_t ld, double2 alpha_,
                                                         double2 beta_, int beta_zero) {
    using T = cplx<R>;
    using V = typename mf<R>::v4;
    constexpr int BLK = 16 * D2;     // elements per block
    constexpr int NT2 = D2 / 16, KS2 = D2 / 4;
    constexpr int NT3 = D3 / 16, KS3 = D3 / 4;
    constexpr int NB = D3;           // blocks per tile
    constexpr int D = BLK * NB;      // elements per tile
    constexpr int WPB = REG_NT / 64;
    constexpr int BPW = NB / WPB;    // blocks per wave (NB is 16 or 32)
    constexpr int LDA2 = D2 + 1, LDA3 = D3 + 1;
    static_assert(NB % WPB == 0, "blocks per tile must be a multiple of the waves");
    extern __shared__ unsigned char lds_raw[];
    T *A2s = (T *)lds_raw;
    T *A3s = A2s + D2 * LDA2;
    constexpr int Dp = D + (D >> 4) + 1;
    R *yre = (R *)(A3s + D3 * LDA3);
    R *yim = yre + Dp;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const T alpha = mk((R)alpha_.x, (R)alpha_.y), beta = mk((R)beta_.x, (R)beta_.y);

    for (int e = tid; e < D2 * D2; e += REG_NT) A2s[(e % D2) + (e / D2) * LDA2] = A2[e];
    for (int e = tid; e < D3 * D3; e += REG_NT) A3s[(e % D3) + (e / D3) * LDA3] = A3[e];
    __syncthreads();  // A2 and A3 staged

    auto tile_base = [&](int64_t c) -> int64_t { return 16 * (c % nrb) + ld * (int64_t)D2 * NB * (c / nrb); };
    // the wave's blocks b = wave + WPB q: mode-2 B-fragments, k-step ks
    // holding (r = lane % 16, i2 = 4 ks + lane / 16)
    T x[BPW][KS2];
    auto load_tile = [&](int64_t base) {
#pragma unroll
        for (int q = 0; q < BPW; ++q)
#pragma unroll
            for (int ks = 0; ks < KS2; ++ks)
                x[q][ks] = in[base + lr + ld * (int64_t)(4 * ks + lk + D2 * (wave + WPB * q))];
    };
    // mode 2 of block q: tile u of D2 = A2 M (rows i2' = 16 u + mf<R>::row(lane, r), columns r = lane % 16)
    auto mode2 = [&](const T *F2, int q, V (&d2r)[NT2], V (&d2i)[NT2]) {
#pragma unroll
        for (int u = 0; u < NT2; ++u) {
            d2r[u] = V{0, 0, 0, 0};
            d2i[u] = V{0, 0, 0, 0};
        }
#pragma unroll
        for (int ks = 0; ks < KS2; ++ks)
#pragma unroll
            for (int u = 0; u < NT2; ++u) cmma4<R>(F2[(16 * u + lr) + (4 * ks + lk) * LDA2], x[q][ks], d2r[u], d2i[u]);
    };
    auto put = [&](int blk, const V (&d2r)[NT2], const V (&d2i)[NT2]) {
#pragma unroll
        for (int u = 0; u < NT2; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = blk * BLK + lr + 16 * (16 * u + mf<R>::row(lane, r));
                yre[pad16r(e)] = d2r[u][r];
                yim[pad16r(e)] = d2i[u][r];
            }
    };
    int64_t c = blockIdx.x;
    if (c < ntiles) load_tile(tile_base(c));
    for (; c < ntiles; c += gridDim.x) {
        const int64_t base = tile_base(c);
        const int64_t nc = c + gridDim.x;
        // the factors stay in LDS: an offset the compiler cannot see through
        // keeps their reads inside the loop instead of hoisting them into
        // registers for the whole launch
        int fo = 0;
        asm volatile("" : "+s"(fo));
        const T *F2 = A2s + fo, *F3 = A3s + fo;
        {
            // the first block is formed while other waves may still read the
            // previous tile's planes, and written after the barrier
            V d2r[NT2], d2i[NT2];
            mode2(F2, 0, d2r, d2i);
            __syncthreads();  // every wave is done reading the previous tile's planes
            put(wave, d2r, d2i);
#pragma unroll
            for (int q = 1; q < BPW; ++q) {
                mode2(F2, q, d2r, d2i);
                put(wave + WPB * q, d2r, d2i);
            }
        }
        __syncthreads();  // the tile's planes are complete
        if (nc < ntiles) load_tile(tile_base(nc));
        // mode 3 over groups of 16 plane elements e = r + 16 i2 (r = lane % 16).
        // A compile-time group count per wave: the wait for the next tile's
        // loads (issued above, before these stores) then counts every store
        // of this mode, where a run-time trip count made the next tile's
        // first MFMAs wait for all but one group's stores as well.
        static_assert((BLK / 16) % WPB == 0, "mode-3 groups per wave");
#pragma unroll
        for (int q = 0; q < BLK / 16 / WPB; ++q) {
            __builtin_amdgcn_sched_barrier(0);  // one group's registers live at a time
            const int g = wave + q * WPB;
            const int e0 = 16 * g;
            V zr[NT3], zi[NT3];
#pragma unroll
            for (int v = 0; v < NT3; ++v) {
                zr[v] = V{0, 0, 0, 0};
                zi[v] = V{0, 0, 0, 0};
            }
            // two k-steps at a time: the next tile's loads are in flight in
            // registers meanwhile
#pragma unroll 2
            for (int ks = 0; ks < KS3; ++ks) {
                const int idx = pad16r(e0 + lr + BLK * (4 * ks + lk));
                const T yv = mk(yre[idx], yim[idx]);
#pragma unroll
                for (int v = 0; v < NT3; ++v) cmma4<R>(F3[(16 * v + lr) + (4 * ks + lk) * LDA3], yv, zr[v], zi[v]);
            }
#pragma unroll
            for (int v = 0; v < NT3; ++v)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    T *o = out + base + lr + ld * (int64_t)(g + D2 * (16 * v + mf<R>::row(lane, r)));
                    T val = cmul(alpha, mk((R)zr[v][r], (R)zi[v][r]));
                    if (!beta_zero) val = cfma(beta, *o, val);
                    st_stream(o, val);
                }
        }
    }
}

}  // namespace

int g_fused_reg = 1;  // 1: this kernel where the factors' sites and sizes fit it

size_t lt_reg_lds(qob_dtype dt, int d2, int D) {
    const size_t r = dt == QOB_C64 ? 8 : 4;
    const size_t dp = (size_t)(D + (D >> 4) + 1);
    return (size_t)d2 * (d2 + 1) * 2 * r + 2 * dp * r;
}

// Launch when the factors sit on sites 1 and 2 (and 3, the last site) with
// d1 = 16, d2 in {16, 32} and d3 in {16, 32} (not both 32); returns
// hipErrorNotSupported otherwise (the caller uses the fused column kernel).
hipError_t launch_lt_reg(qob_dtype dt, const FusedModes &M, const void *in, void *out, int64_t ncols, double2 alpha,
                         double2 beta, hipStream_t s) {
    if (!g_fused_reg || M.nmodes < 2 || M.nmodes > 3) return hipErrorNotSupported;
    if (M.n1[0] != 1 || M.d[0] != 16 || M.n1[1] != 16 || (M.d[1] != 16 && M.d[1] != 32)) return hipErrorNotSupported;
    const int d2 = M.d[1], blk = 16 * d2;
    if (M.D % blk) return hipErrorNotSupported;
    const int nblk = M.D / blk;
    int d3 = 0;
    if (M.nmodes == 3) {
        d3 = M.d[2];
        if (M.n1[2] != blk || (d3 != 16 && d3 != 32) || d3 != nblk) return hipErrorNotSupported;
    }
    const size_t lds = d3 ? lt_reg_lds(dt, d2, M.D) : (size_t)d2 * (d2 + 1) * (dt == QOB_C64 ? 16 : 8);
    if (lds > 160 * 1024) return hipErrorNotSupported;
    if (ncols == 0) return hipSuccess;
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    const int ncu = device_cu_count();
#define QOB_LR(RR, DD2, DD3)                                                                                    \
    do {                                                                                                        \
        auto k = k_lt_reg<RR, DD2, DD3>;                                                                        \
        static std::atomic<int> nres{0};                                                                        \
        if (!nres.load()) {                                                                                     \
            hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,     \
                                               160 * 1024);                                                     \
            if (e != hipSuccess) return e;                                                                      \
            int nb = 0;                                                                                         \
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, REG_NT, lds);                              \
            if (e != hipSuccess) return e;                                                                      \
            nres.store(nb > 0 ? nb : 1);                                                                        \
        }                                                                                                       \
        const unsigned g = (unsigned)std::min<int64_t>(ncols, (int64_t)ncu * nres.load());                      \
        hipLaunchKernelGGL(k, dim3(g), dim3(REG_NT), lds, s, (const cplx<RR> *)M.A[0], (const cplx<RR> *)M.A[1], \
                           (const cplx<RR> *)(d3 ? M.A[2] : nullptr), (const cplx<RR> *)in, (cplx<RR> *)out,    \
                           ncols, nblk, alpha, beta, bz);                                                       \
        return hipGetLastError();                                                                               \
    } while (0)
    // (d2, d3) = (32, 32) is left to the fused column kernel: its eight
    // mode-3 fragment pairs per lane do not fit next to the rest
    if (dt == QOB_C64) {
        if (d2 == 16 && d3 == 0) QOB_LR(double, 16, 0);
        if (d2 == 16 && d3 == 16) QOB_LR(double, 16, 16);
        if (d2 == 16 && d3 == 32) QOB_LR(double, 16, 32);
        if (d2 == 32 && d3 == 0) QOB_LR(double, 32, 0);
        if (d2 == 32 && d3 == 16) QOB_LR(double, 32, 16);
    } else {
        if (d2 == 16 && d3 == 0) QOB_LR(float, 16, 0);
        if (d2 == 16 && d3 == 16) QOB_LR(float, 16, 16);
        if (d2 == 16 && d3 == 32) QOB_LR(float, 16, 32);
        if (d2 == 32 && d3 == 0) QOB_LR(float, 32, 0);
        if (d2 == 32 && d3 == 16) QOB_LR(float, 32, 16);
    }
#undef QOB_LR
    return hipErrorNotSupported;
}

// Right application in 16-row tiles (see k_lt_reg_tile): the tile's two
// factors sit on its sites 2 and 3 (FusedModes n1 = 16 and 16 d2, the
// tile exactly 16 x d2 x d3 elements), d2 and d3 in {16, 32}, not both 32.
// Returns hipErrorNotSupported otherwise (the caller runs the fused column
// kernel on the tiles).
hipError_t launch_lt_reg_tile(qob_dtype dt, const FusedModes &M, const void *in, void *out, int64_t ntiles,
                              int64_t nrb, int64_t ld, double2 alpha, double2 beta, hipStream_t s) {
    if (!g_fused_reg || M.nmodes != 2 || M.n1[0] != 16) return hipErrorNotSupported;
    const int d2 = M.d[0], d3 = M.d[1];
    if ((d2 != 16 && d2 != 32) || (d3 != 16 && d3 != 32) || (d2 == 32 && d3 == 32)) return hipErrorNotSupported;
    if (M.n1[1] != 16 * d2 || M.D != 16 * d2 * d3) return hipErrorNotSupported;
    if (nrb < 1 || ld < 16 * nrb || ntiles < 0 || ntiles % nrb) return hipErrorInvalidValue;
    const size_t lds = lt_reg_lds(dt, d2, M.D) + (size_t)d3 * (d3 + 1) * (dt == QOB_C64 ? 16 : 8);  // + A3
    if (lds > 160 * 1024) return hipErrorNotSupported;
    if (ntiles == 0) return hipSuccess;
    const int bz = (beta.x == 0.0 && beta.y == 0.0);
    const int ncu = device_cu_count();
#define QOB_LT(RR, DD2, DD3)                                                                                    \
    do {                                                                                                        \
        auto k = k_lt_reg_tile<RR, DD2, DD3>;                                                                   \
        static std::atomic<int> nres{0};                                                                        \
        if (!nres.load()) {                                                                                     \
            hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,     \
                                               160 * 1024);                                                     \
            if (e != hipSuccess) return e;                                                                      \
            int nb = 0;                                                                                         \
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, REG_NT, lds);                              \
            if (e != hipSuccess) return e;                                                                      \
            nres.store(nb > 0 ? nb : 1);                                                                        \
        }                                                                                                       \
        const unsigned g = (unsigned)std::min<int64_t>(ntiles, (int64_t)ncu * nres.load());                     \
        hipLaunchKernelGGL(k, dim3(g), dim3(REG_NT), lds, s, (const cplx<RR> *)M.A[0], (const cplx<RR> *)M.A[1], \
                           (const cplx<RR> *)in, (cplx<RR> *)out, ntiles, nrb, ld, alpha, beta, bz);            \
        return hipGetLastError();                                                                               \
    } while (0)
    if (dt == QOB_C64) {
        if (d2 == 16 && d3 == 16) QOB_LT(double, 16, 16);
        if (d2 == 16 && d3 == 32) QOB_LT(double, 16, 32);
        if (d2 == 32 && d3 == 16) QOB_LT(double, 32, 16);
    } else {
        if (d2 == 16 && d3 == 16) QOB_LT(float, 16, 16);
        if (d2 == 16 && d3 == 32) QOB_LT(float, 16, 32);
        if (d2 == 32 && d3 == 16) QOB_LT(float, 32, 16);
    }
#undef QOB_LT
    return hipErrorNotSupported;
}

}  // namespace qob
