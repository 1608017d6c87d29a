// Fused LazySum of qubit Pauli-type terms, for one state or a batch of
// states of different lengths in one persistent launch.
//
// Reference: mul!(::Ket, ::LazySum, ::Ket) (src/operators_lazysum.jl:189-200)
// runs each LazyTensor term through the recursive scatter
// _gemm_recursive_lazy_dense (src/operators_lazytensor.jl:652-685), one full
// read-modify-write of the result per term.  When every site is a qubit and
// every factor is monomial (at most one entry per row: sigma_x/y/z, sigma_+-,
// number, projectors), a term maps basis state J to J ^ F_t with weight
// c_t * prod_s v_s[bit_s(J)].  The host groups the terms by flip mask F and
// folds each group's weight into either a constant or a sum of small tables
// indexed by a window of <= 8 bits of J:
//   out[J] = alpha * sum_F w_F(J) * in[J ^ F] + gamma * in[J] + beta * out[J]
// gamma fuses the Ket update of a time step (psi + c H psi) into the apply.
//
// A workgroup owns a tile of 2^tb consecutive basis states: the tile is read
// once into LDS, flips inside the tile read LDS, flips of higher bits read
// global memory (coalesced, served by L2 / Infinity Cache).  Work items
// (state, tile) of all states of a batch share one launch.
//
// Program of a state (int32, read with wave-uniform scalar loads; records
// are multiples of 4 ints, so every record starts 16-byte aligned):
//   header            [nconst, nwgroups, 0, 0]
//   nconst records    [F, 0, 0, 0, w.re, w.im]   (w as two doubles = 4 ints)
//                     flip F with a constant, nonzero weight w
//   nwgroups records  [F, nwin, 0, 0, c.re, c.im] each followed by nwin
//                     windows [shift, nbits, table_off, 0]; the weight is
//                     w_F(J) = c + sum_win vals[table_off + ((J >> shift) & (2^nbits - 1))]
// Constant records have a fixed stride: the record of group g + 1 does not
// depend on group g and is loaded while group g is applied.
#include "qob_internal.h"

#include <algorithm>
#include <atomic>

namespace qob {

namespace {

// logical workgroup id such that each of the 8 XCDs owns a contiguous range
// (linear id b runs on XCD b % 8): tiles of one state that differ in high
// bits are then processed on the same XCD and share its L2.
__device__ __forceinline__ uint32_t qb_xcd_remap(uint32_t b, uint32_t G) {
    const uint32_t x = b & 7u, i = b >> 3;
    const uint32_t q = G >> 3, r = G & 7u;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// The items one workgroup walks.  Per-XCD ranges (flags & 8): workgroup b
// runs on XCD x = b % 8 and walks [off[x], off[x + 1]) with a stride of the
// number of workgroups on that XCD (the grid has at least 8 workgroups then).
// Otherwise rounds of G items over the whole list, each XCD a contiguous run
// of every round when flags & 4.
struct ItemWalk {
    int it, end, step;
};
__device__ __forceinline__ ItemWalk qb_walk(int flags, const QubitSplit &xs, int nitems) {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    if (flags & 8) {
        const uint32_t x = b & 7u;
        return {xs.off[x] + (int)(b >> 3), xs.off[x + 1], (int)((G - x + 7u) >> 3)};
    }
    const uint32_t lb = (flags & 4) ? qb_xcd_remap(b, G) : b;
    return {(int)lb, nitems, (int)G};
}

template <typename T> __device__ __forceinline__ T to_t(double2 a) {
    T r; r.x = (typename real_of<T>::type)a.x; r.y = (typename real_of<T>::type)a.y; return r;
}

// complex weight stored as two doubles in program words r[0..3]
template <typename T> __device__ __forceinline__ T rec_weight(const int32_t *r) {
    T w;
    w.x = (typename real_of<T>::type)__hiloint2double(r[1], r[0]);
    w.y = (typename real_of<T>::type)__hiloint2double(r[3], r[2]);
    return w;
}
// the same, held in scalar registers: a weight carried from one loop
// iteration to the next is otherwise re-read with a vector load, whose wait
// would also wait for every partner load issued before it
template <typename T> __device__ __forceinline__ T rec_weight_s(const int32_t *r) {
    T w;
    w.x = (typename real_of<T>::type)__hiloint2double(__builtin_amdgcn_readfirstlane(r[1]),
                                                      __builtin_amdgcn_readfirstlane(r[0]));
    w.y = (typename real_of<T>::type)__hiloint2double(__builtin_amdgcn_readfirstlane(r[3]),
                                                      __builtin_amdgcn_readfirstlane(r[2]));
    return w;
}

// acc += w * x for every element of the lane, w uniform; a real weight
// needs two FMAs instead of four.  Partner of element k (tile index
// e = tid + k NTH) under flip F is e ^ F = (tid ^ F) ^ (k NTH): one XOR with a
// compile-time constant per element, on the byte offset tpb = (tid ^ F) * sizeof(T).
// NT: a partner tile read once (non-temporal load: it is not kept in L2 at
// the expense of tiles read again soon)
template <bool NT, typename T> __device__ __forceinline__ T qb_ld(const T *p) {
    if constexpr (NT) return ldg_nt(p);
    else return ldg(p);
}

// acc[k] += w * x[k] with w uniform, updated in place: the w.y terms first
// (skipped for a real weight), then the w.x terms, which is cfma's order, so
// both cases give cfma's bits (a real weight never multiplies a NaN
// imaginary partner part).  Both branches write the accumulators where they
// are, so the compiler needs no copies to merge them.
template <typename T, int EPT>
__device__ __forceinline__ void qb_acc(T wt, const T (&x)[EPT], T (&acc)[EPT]) {
    if (wt.y != 0) {
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            acc[k].x = fma(-wt.y, x[k].y, acc[k].x);
            acc[k].y = fma(wt.y, x[k].x, acc[k].y);
        }
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        acc[k].x = fma(wt.x, x[k].x, acc[k].x);
        acc[k].y = fma(wt.x, x[k].y, acc[k].y);
    }
}

// GL: basep points to global memory (a partner tile), else to the LDS tile
template <typename T, int NTH, int EPT, bool GL = false, bool NT = false>
__device__ __forceinline__ void qb_uniform(T wt, const char *basep, uint32_t tpb, T (&acc)[EPT]) {
    T x[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const T *p = (const T *)(basep + (tpb ^ (uint32_t)(k * NTH * (int)sizeof(T))));
        if constexpr (GL) x[k] = qb_ld<NT>(p);
        else x[k] = *p;
    }
    qb_acc<T, EPT>(wt, x, acc);
}

// acc += w * x[k ^ f] for every element k of the lane: a flip of tile bits
// at or above log2(NTH) (F = f NTH) pairs elements of one lane, read from
// the lane's registers; f = 0 is the diagonal.  f is wave-uniform, so the
// branch per value keeps every register index a compile-time constant.
template <typename T, int EPT>
__device__ __forceinline__ void qb_regflip(T wt, const T (&xv)[EPT], uint32_t f, T (&acc)[EPT]) {
#pragma unroll
    for (int ff = 0; ff < EPT; ++ff) {
        if (f == (uint32_t)ff) {
            T x[EPT];
#pragma unroll
            for (int k = 0; k < EPT; ++k) x[k] = xv[k ^ ff];
            qb_acc<T, EPT>(wt, x, acc);
        }
    }
}

// the same with a per-element weight; a zero weight selects (a NaN partner
// must not propagate)
template <typename T, int EPT>
__device__ __forceinline__ void qb_regflip_w(const T (&wt)[EPT], const T (&xv)[EPT], uint32_t f, T (&acc)[EPT]) {
#pragma unroll
    for (int ff = 0; ff < EPT; ++ff) {
        if (f == (uint32_t)ff) {
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const T v = cfma(wt[k], xv[k ^ ff], acc[k]);
                const bool nz = wt[k].x != 0 || wt[k].y != 0;
                acc[k].x = nz ? v.x : acc[k].x;
                acc[k].y = nz ? v.y : acc[k].y;
            }
        }
    }
}

// Tile coordinates of the first pass: tile element e is basis state base + e.
struct LoMap {
    uint32_t base, lowmask;
    int tb;
    __device__ __forceinline__ uint32_t J(uint32_t e) const { return base + e; }
    __device__ __forceinline__ bool in_tile(uint32_t F) const { return (F >> tb) == 0; }
    __device__ __forceinline__ uint32_t fq(uint32_t F) const { return F; }
};

// Tile coordinates of the second pass (strided tile, see k_qubit_hi).
template <int TBITS> struct HiMap {
    uint32_t mid, lowmask;  // lowmask = 2^c - 1
    int c;
    __device__ __forceinline__ uint32_t J(uint32_t q) const {
        return ((q >> c) << TBITS) | mid | (q & lowmask);
    }
    __device__ __forceinline__ bool in_tile(uint32_t F) const {
        return ((F & ((1u << TBITS) - 1u)) >> c) == 0;
    }
    __device__ __forceinline__ uint32_t fq(uint32_t F) const {
        return ((F >> TBITS) << c) | (F & ((1u << TBITS) - 1u));
    }
};

// acc[k] += w_F(J) * in[J ^ F] over the groups of one program, for the
// EPT elements e = tid + k NTH of the lane's tile.  Elements past a partial
// tile compute on unused LDS words and are never stored; partners of a
// partial tile never leave it (the state is a single tile then).
template <typename T, int NTH, int EPT, bool LO, class Map>
__device__ __forceinline__ void qb_run(const int32_t *__restrict__ pg, const T *__restrict__ vals, const T *tile,
                                       const T *__restrict__ in, const Map &m, const T (&xv)[EPT], T (&acc)[EPT],
                                       uint32_t far_tiles = 0xffffffffu) {
    const uint32_t tid = threadIdx.x;
    const int nconst = pg[0];
    const int nwg = pg[1];
    const int32_t *r = pg + 4;
    // register flips in ComplexF32 only: the lane's copy of its elements
    // costs ComplexF64 16 VGPRs and 5% (measured), ComplexF32 gains 1.5%
    constexpr bool REG = sizeof(T) == 8;
    if constexpr (LO) {
        // Partner values of group gr into dst, from global memory: the tile
        // of basis states (base + e) ^ F = (base ^ Fhigh) + (e ^ Flow); also
        // right (if slower than LDS) for a flip inside the tile.
        auto issue = [&](uint32_t Fp, T (&dst)[EPT]) {
            const char *gp = (const char *)(in + (m.J(0) ^ (Fp & ~m.lowmask)));
            const uint32_t tpb = (tid ^ (Fp & m.lowmask)) * (uint32_t)sizeof(T);
            if ((Fp >> m.tb) >= far_tiles) {
#pragma unroll
                for (int k = 0; k < EPT; ++k)
                    dst[k] = ldg_nt((const T *)(gp + (tpb ^ (uint32_t)(k * NTH * (int)sizeof(T)))));
            } else {
#pragma unroll
                for (int k = 0; k < EPT; ++k)
                    dst[k] = ldg((const T *)(gp + (tpb ^ (uint32_t)(k * NTH * (int)sizeof(T)))));
            }
        };
        auto accum = [&](T wt, const T (&x)[EPT]) { qb_acc<T, EPT>(wt, x, acc); };
        // Constant groups are in ascending F: groups [0, nin) flip bits inside
        // the tile (header word 2), the rest read partner tiles.  Partner
        // loads run one group ahead through two buffers, the first group's
        // during the in-tile groups; accumulation stays in program order.
        // (nin is computed on the host for this tile size; the tile mask keeps
        // every LDS read inside the tile whatever the program holds)
        const int nin = min(max(pg[2], 0), nconst);
        T A[EPT], B[EPT];
        // the first out-of-tile group's loads fly during the in-tile groups
        if (nin < nconst) issue((uint32_t)__builtin_amdgcn_readfirstlane(r[8 * nin]), A);
        if (nin > 0) {
            uint32_t F = (uint32_t)__builtin_amdgcn_readfirstlane(r[0]);
            T wt = rec_weight_s<T>(r + 4);
            for (int g = 0; g < nin; ++g) {
                const int32_t *rn = r + 8 * (g + 1 < nin ? g + 1 : g);
                const uint32_t Fn = (uint32_t)__builtin_amdgcn_readfirstlane(rn[0]);
                const T wn = rec_weight_s<T>(rn + 4);
                if (REG && (F & (uint32_t)(NTH - 1)) == 0)  // flips of register bits only
                    qb_regflip<T, EPT>(wt, xv, (F & m.lowmask) / (uint32_t)NTH, acc);
                else
                    qb_uniform<T, NTH, EPT>(wt, (const char *)tile,
                                            ((tid ^ m.fq(F)) & m.lowmask) * (uint32_t)sizeof(T), acc);
                F = Fn;
                wt = wn;
            }
        }
        if (nin < nconst) {
            // Two buffers in a loop over pairs of groups, every group's loads
            // issued unconditionally (the last one's once more when the count
            // is even), so that each wait counts the same loads on every
            // path: A holds group g while B fills with group g + 1.
            int g = nin;
            const int last = nconst - 1;
            const int npairs = (nconst - nin) >> 1;
            for (int pq = 0; pq < npairs; ++pq) {
                const uint32_t F1 = (uint32_t)__builtin_amdgcn_readfirstlane(r[8 * (g + 1)]);
                const uint32_t F2 = (uint32_t)__builtin_amdgcn_readfirstlane(r[8 * min(g + 2, last)]);
                const T w1 = rec_weight_s<T>(r + 8 * g + 4);
                const T w2 = rec_weight_s<T>(r + 8 * (g + 1) + 4);
                issue(F1, B);
                accum(w1, A);
                issue(F2, A);
                accum(w2, B);
                g += 2;
            }
            if (g < nconst) accum(rec_weight_s<T>(r + 8 * g + 4), A);
        }
    } else if (nconst > 0) {
        uint32_t F = (uint32_t)r[0];
        T wt = rec_weight<T>(r + 4);
        for (int g = 0; g < nconst; ++g) {
            const int32_t *rn = r + 8 * (g + 1 < nconst ? g + 1 : g);
            const uint32_t Fn = (uint32_t)rn[0];
            const T wn = rec_weight<T>(rn + 4);
            if (m.in_tile(F)) {
                qb_uniform<T, NTH, EPT>(wt, (const char *)tile, (tid ^ m.fq(F)) * (uint32_t)sizeof(T), acc);
            } else {
#pragma unroll
                for (int k = 0; k < EPT; ++k) acc[k] = cfma(wt, ldg(in + (m.J(tid + k * NTH) ^ F)), acc[k]);
            }
            F = Fn;
            wt = wn;
        }
    }
    r += 8 * nconst;
    for (int g = 0; g < nwg; ++g) {
        const uint32_t F = (uint32_t)r[0];
        const int nwin = r[1];
        const T c0 = rec_weight<T>(r + 4);
        r += 8;
        T wt[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) wt[k] = c0;
        for (int q = 0; q < nwin; ++q, r += 4) {
            const int sh = r[0], toff = r[2];
            const uint32_t mask = (1u << r[1]) - 1u;
            bool above = false;
            if constexpr (LO) above = sh >= m.tb;
            if (above) {
                // window above the tile: one entry for the whole tile
                const T v = vals[toff + (int)((m.J(0) >> sh) & mask)];
#pragma unroll
                for (int k = 0; k < EPT; ++k) wt[k] = cadd(wt[k], v);
            } else if (LO && (NTH >> sh) > (int)mask) {
                // window inside the lane index bits: one entry per lane
                const T v = vals[toff + (int)((tid >> sh) & mask)];
#pragma unroll
                for (int k = 0; k < EPT; ++k) wt[k] = cadd(wt[k], v);
            } else {
#pragma unroll
                for (int k = 0; k < EPT; ++k)
                    wt[k] = cadd(wt[k], vals[toff + (int)((m.J(tid + k * NTH) >> sh) & mask)]);
            }
        }
        // a zero weight must not propagate a NaN / Inf partner (rows absent
        // from the factors): select instead of multiply
        if (REG && LO && m.in_tile(F) && (F & (uint32_t)(NTH - 1)) == 0) {
            qb_regflip_w<T, EPT>(wt, xv, (F & m.lowmask) / (uint32_t)NTH, acc);
        } else if (m.in_tile(F)) {
            const uint32_t tpb = (tid ^ m.fq(F)) * (uint32_t)sizeof(T);
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const T x = *(const T *)((const char *)tile + (tpb ^ (uint32_t)(k * NTH * (int)sizeof(T))));
                const T v = cfma(wt[k], x, acc[k]);
                const bool nz = wt[k].x != 0 || wt[k].y != 0;
                acc[k].x = nz ? v.x : acc[k].x;
                acc[k].y = nz ? v.y : acc[k].y;
            }
        } else {
#pragma unroll
            for (int k = 0; k < EPT; ++k) {
                const T x = ldg(in + (m.J(tid + k * NTH) ^ F));
                const T v = cfma(wt[k], x, acc[k]);
                const bool nz = wt[k].x != 0 || wt[k].y != 0;
                acc[k].x = nz ? v.x : acc[k].x;
                acc[k].y = nz ? v.y : acc[k].y;
            }
        }
    }
}

// TBITS: log2 of the tile; each lane owns EPT = 2^TBITS / NTH elements of a
// tile (e = tid + k * NTH) and the group loop is outermost, so the program is
// decoded once per group and the lane's EPT partner loads are independent.
template <typename T, int NTH, int TBITS>
__global__ void __launch_bounds__(NTH) k_qubit(const QubitStateDev *__restrict__ states,
                                               const QubitItemDev *__restrict__ items, int nitems,
                                               const int32_t *__restrict__ prog, const T *__restrict__ vals,
                                               double2 alpha_, double2 beta_, double2 gamma_, int flags,
                                               QubitStateDev single, int ntiles1, QubitSplit xs) {
    constexpr int EPT = (1 << TBITS) / NTH;
    extern __shared__ unsigned char lds_raw[];
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_), gamma = to_t<T>(gamma_);
    const bool beta_zero = flags & 1, gamma_zero = flags & 2;
    // flags & 256: two tile buffers in LDS, used by alternate items.  A
    // buffer is rewritten two items after its last reads, and every wave has
    // passed the barrier after the last item's store since then, so one
    // barrier per item suffices.
    const bool dbuf = flags & 256;
    int par = 0;
    const int tid = threadIdx.x;
    // per-XCD ranges keep the partner tiles of a state's high-bit flips in the
    // L2 of the XCD that reads them; rounds of G keep all workgroups on
    // neighbouring tiles at the same time
    const ItemWalk wk = qb_walk(flags, xs, nitems);
    // with flags & 16, partner tiles at least as many tiles away as an XCD
    // has workgroups lie outside the run of tiles in flight there: read once
    // (flags >> 5: the distance is that run's length halved that many times)
    const uint32_t far_tiles =
        !(flags & 16) ? 0xffffffffu
                      : max(1u, ((flags & 8) ? (uint32_t)wk.step : (uint32_t)gridDim.x >> 3) >> ((flags >> 5) & 7));
    for (int it = wk.it; it < wk.end; it += wk.step) {
        QubitItemDev w;
        QubitStateDev S;
        if (states) {
            w = items[it];
            S = states[w.state];
        } else {
            // columns form: state c is column c of a 2^nbits x ncols matrix
            const int c = it / ntiles1;
            w.state = c;
            w.tile = it - c * ntiles1;
            S = single;
            S.in = (const T *)single.in + ((int64_t)c << single.nbits);
            S.out = (T *)single.out + ((int64_t)c << single.nbits);
        }
        const int tb = S.nbits < TBITS ? S.nbits : TBITS;
        const int tsize = 1 << tb;
        const uint32_t base = (uint32_t)w.tile << tb;
        const T *__restrict__ in = (const T *)S.in;
        T *__restrict__ out = (T *)S.out;
        // a second pass already wrote alpha * high + beta * out: add to it
        const bool add_out = S.prog_hi_off >= 0;
        const bool bz = beta_zero && !add_out;
        const T beff = add_out ? mk((typename real_of<T>::type)1, (typename real_of<T>::type)0) : beta;
        // the lane keeps its own elements in registers as well: flips of
        // the register bits (e = tid + k NTH, k = 0..EPT-1), the diagonal
        // and the gamma term read them there
        T xv[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = tid + k * NTH;
            xv[k] = e < tsize ? ldg(in + base + e) : czero<T>();
        }
        T *tile = (T *)lds_raw + (par << TBITS);
        par ^= (int)dbuf;
        if (!dbuf) __syncthreads();  // the previous item's tile reads are finished
#pragma unroll
        for (int k = 0; k < EPT; ++k) tile[tid + k * NTH] = xv[k];
        __syncthreads();
        T acc[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) acc[k] = czero<T>();
        const LoMap m{base, (uint32_t)tsize - 1u, tb};
        qb_run<T, NTH, EPT, true>(prog + S.prog_off, vals, tile, in, m, xv, acc, far_tiles);
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = tid + k * NTH;
            if (e < tsize) {
                const uint32_t J = base + (uint32_t)e;
                T r = cmul(alpha, acc[k]);
                if (!gamma_zero) r = cfma(gamma, sizeof(T) == 8 ? xv[k] : tile[e], r);
                if (!bz) r = cfma(beff, ldg(out + J), r);
                stg_stream(out + J, r);
            }
        }
    }
}

// ComplexF32 in pairs of adjacent elements.  Lane tid holds the pairs
// P = tid + j NTH (j < EPT / 2) of its tile, elements 2P (register 2j) and
// 2P + 1 (register 2j + 1), so every global and LDS access moves 16 bytes:
// ComplexF32 elements one at a time are 8-byte accesses, twice the vector
// memory instructions per byte, and the texture addressers bound the
// kernel.  A flip F takes pair P to pair P ^ (F >> 1), halves swapped when F
// is odd.  Every element receives the same terms in the same order through
// the same operations (qb_acc, cfma with zero select) as in k_qubit, so the
// results are the same bits.
typedef float qb_f4 __attribute__((ext_vector_type(4)));
constexpr int qb_lg(int n) { return n <= 1 ? 0 : 1 + qb_lg(n / 2); }

template <bool GL>
__device__ __forceinline__ void qp_load(const char *p, float2 &lo, float2 &hi) {
    qb_f4 v;
    if constexpr (GL) v = *(const QOB_GLOBAL qb_f4 *)p;
    else v = *(const qb_f4 *)p;
    lo.x = v.x; lo.y = v.y; hi.x = v.z; hi.y = v.w;
}

// x[2j + h] <- pair at byte offset tpb ^ (j NTH 16) from basep
template <int NTH, int EPT, bool GL>
__device__ __forceinline__ void qp_load_all(const char *basep, uint32_t tpb, float2 (&x)[EPT]) {
#pragma unroll
    for (int j = 0; j < EPT / 2; ++j)
        qp_load<GL>(basep + (tpb ^ (uint32_t)(j * NTH * 16)), x[2 * j], x[2 * j + 1]);
}

// acc += w * x, halves swapped when sw (a uniform branch; the swap is a
// register renaming)
template <int EPT>
__device__ __forceinline__ void qp_acc(float2 wt, const float2 (&x)[EPT], bool sw, float2 (&acc)[EPT]) {
    if (!sw) {
        qb_acc<float2, EPT>(wt, x, acc);
    } else {
        float2 xs[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) xs[k] = x[k ^ 1];
        qb_acc<float2, EPT>(wt, xs, acc);
    }
}

// the same with per-element weights and a zero weight selecting
template <int EPT>
__device__ __forceinline__ void qp_acc_w(const float2 (&wt)[EPT], const float2 (&x)[EPT], bool sw,
                                         float2 (&acc)[EPT]) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const float2 v = cfma(wt[k], sw ? x[k ^ 1] : x[k], acc[k]);
        const bool nz = wt[k].x != 0 || wt[k].y != 0;
        acc[k].x = nz ? v.x : acc[k].x;
        acc[k].y = nz ? v.y : acc[k].y;
    }
}

// a flip of bit 0 and pair-register bits only (f = pair flip / NTH): the
// partners are registers of the lane's own elements.  The accumulation sits
// in each branch so that every register index stays a compile-time constant
// (a branch that only copied would be merged into an indexed read from
// memory).
template <int EPT>
__device__ __forceinline__ void qp_regflip(float2 wt, const float2 (&xv)[EPT], uint32_t f, bool sw,
                                           float2 (&acc)[EPT]) {
#pragma unroll
    for (int ff = 0; ff < EPT / 2; ++ff) {
        if (f == (uint32_t)ff) {
            float2 x[EPT];
#pragma unroll
            for (int k = 0; k < EPT; ++k) x[k] = xv[k ^ (2 * ff)];
            qp_acc<EPT>(wt, x, sw, acc);
        }
    }
}
template <int EPT>
__device__ __forceinline__ void qp_regflip_w(const float2 (&wt)[EPT], const float2 (&xv)[EPT], uint32_t f, bool sw,
                                             float2 (&acc)[EPT]) {
#pragma unroll
    for (int ff = 0; ff < EPT / 2; ++ff) {
        if (f == (uint32_t)ff) {
            float2 x[EPT];
#pragma unroll
            for (int k = 0; k < EPT; ++k) x[k] = xv[k ^ (2 * ff)];
            qp_acc_w<EPT>(wt, x, sw, acc);
        }
    }
}

template <int NTH, int EPT>
__device__ __forceinline__ void qb_run_pair(const int32_t *__restrict__ pg, const float2 *__restrict__ vals,
                                            const float2 *tile, const float2 *__restrict__ in, uint32_t base,
                                            uint32_t lowmask, int tb, const float2 (&xv)[EPT], float2 (&acc)[EPT]) {
    using T = float2;
    constexpr int LNTH = qb_lg(NTH);
    const uint32_t tid = threadIdx.x;
    const uint32_t lowp = lowmask >> 1;  // pair index mask of the tile
    const int nconst = pg[0];
    const int nwg = pg[1];
    const int32_t *r = pg + 4;
    // byte offset of the lane's partner pair 0 under flip F inside a tile
    auto tpb_of = [&](uint32_t F) -> uint32_t { return ((tid ^ (F >> 1)) & lowp) * 16u; };
    // partner pairs of an out-of-tile flip, from global memory
    auto issue = [&](uint32_t Fp, T (&dst)[EPT]) {
        const char *gp = (const char *)(in + (base ^ (Fp & ~lowmask)));
        qp_load_all<NTH, EPT, true>(gp, tpb_of(Fp), dst);
    };
    auto accum = [&](T wt, const T (&x)[EPT], uint32_t F) { qp_acc<EPT>(wt, x, (F & 1u) != 0, acc); };
    const int nin = min(max(pg[2], 0), nconst);
    T A[EPT], B[EPT];
    if (nin < nconst) issue((uint32_t)__builtin_amdgcn_readfirstlane(r[8 * nin]), A);
    if (nin > 0) {
        uint32_t F = (uint32_t)__builtin_amdgcn_readfirstlane(r[0]);
        T wt = rec_weight_s<T>(r + 4);
        for (int g = 0; g < nin; ++g) {
            const int32_t *rn = r + 8 * (g + 1 < nin ? g + 1 : g);
            const uint32_t Fn = (uint32_t)__builtin_amdgcn_readfirstlane(rn[0]);
            const T wn = rec_weight_s<T>(rn + 4);
            if (((F >> 1) & (uint32_t)(NTH - 1)) == 0) {  // flips of bit 0 and pair-register bits only
                qp_regflip<EPT>(wt, xv, (F & lowmask) >> (LNTH + 1), (F & 1u) != 0, acc);
            } else {
                T x[EPT];
                qp_load_all<NTH, EPT, false>((const char *)tile, tpb_of(F), x);
                qp_acc<EPT>(wt, x, (F & 1u) != 0, acc);
            }
            F = Fn;
            wt = wn;
        }
    }
    if (nin < nconst) {
        int g = nin;
        const int last = nconst - 1;
        const int npairs = (nconst - nin) >> 1;
        for (int pq = 0; pq < npairs; ++pq) {
            const uint32_t F0 = (uint32_t)__builtin_amdgcn_readfirstlane(r[8 * g]);
            const uint32_t F1 = (uint32_t)__builtin_amdgcn_readfirstlane(r[8 * (g + 1)]);
            const uint32_t F2 = (uint32_t)__builtin_amdgcn_readfirstlane(r[8 * min(g + 2, last)]);
            const T w1 = rec_weight_s<T>(r + 8 * g + 4);
            const T w2 = rec_weight_s<T>(r + 8 * (g + 1) + 4);
            issue(F1, B);
            accum(w1, A, F0);
            issue(F2, A);
            accum(w2, B, F1);
            g += 2;
        }
        if (g < nconst) accum(rec_weight_s<T>(r + 8 * g + 4), A, (uint32_t)__builtin_amdgcn_readfirstlane(r[8 * g]));
    }
    r += 8 * nconst;
    for (int g = 0; g < nwg; ++g) {
        const uint32_t F = (uint32_t)r[0];
        const int nwin = r[1];
        const T c0 = rec_weight<T>(r + 4);
        r += 8;
        T wt[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) wt[k] = c0;
        for (int q = 0; q < nwin; ++q, r += 4) {
            const int sh = r[0], toff = r[2];
            const uint32_t mask = (1u << r[1]) - 1u;
            if (sh >= tb) {
                // window above the tile: one entry for the whole tile
                const T v = vals[toff + (int)((base >> sh) & mask)];
#pragma unroll
                for (int k = 0; k < EPT; ++k) wt[k] = cadd(wt[k], v);
            } else if (sh >= 1 && sh - 1 + r[1] <= LNTH) {
                // window inside the lane's pair index bits: one entry per lane
                const T v = vals[toff + (int)((tid >> (sh - 1)) & mask)];
#pragma unroll
                for (int k = 0; k < EPT; ++k) wt[k] = cadd(wt[k], v);
            } else {
#pragma unroll
                for (int k = 0; k < EPT; ++k) {
                    const uint32_t J = base + 2u * (tid + (uint32_t)((k >> 1) * NTH)) + (uint32_t)(k & 1);
                    wt[k] = cadd(wt[k], vals[toff + (int)((J >> sh) & mask)]);
                }
            }
        }
        if ((F >> tb) == 0 && ((F >> 1) & (uint32_t)(NTH - 1)) == 0) {
            qp_regflip_w<EPT>(wt, xv, (F & lowmask) >> (LNTH + 1), (F & 1u) != 0, acc);
        } else {
            T x[EPT];
            if ((F >> tb) == 0)
                qp_load_all<NTH, EPT, false>((const char *)tile, tpb_of(F), x);
            else
                qp_load_all<NTH, EPT, true>((const char *)(in + (base ^ (F & ~lowmask))), tpb_of(F), x);
            qp_acc_w<EPT>(wt, x, (F & 1u) != 0, acc);
        }
    }
}

template <int NTH, int TBITS>
__global__ void __launch_bounds__(NTH) k_qubit_pair(const QubitStateDev *__restrict__ states,
                                                    const QubitItemDev *__restrict__ items, int nitems,
                                                    const int32_t *__restrict__ prog, const float2 *__restrict__ vals,
                                                    double2 alpha_, double2 beta_, double2 gamma_, int flags,
                                                    QubitStateDev single, int ntiles1, QubitSplit xs) {
    using T = float2;
    constexpr int EPT = (1 << TBITS) / NTH;
    static_assert(EPT >= 2 && EPT % 2 == 0, "pairs per lane");
    extern __shared__ unsigned char lds_raw[];
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_), gamma = to_t<T>(gamma_);
    const bool beta_zero = flags & 1, gamma_zero = flags & 2;
    const bool dbuf = flags & 256;
    int par = 0;
    const int tid = threadIdx.x;
    const ItemWalk wk = qb_walk(flags, xs, nitems);
    for (int it = wk.it; it < wk.end; it += wk.step) {
        QubitItemDev w;
        QubitStateDev S;
        if (states) {
            w = items[it];
            S = states[w.state];
        } else {
            const int c = it / ntiles1;
            w.state = c;
            w.tile = it - c * ntiles1;
            S = single;
            S.in = (const T *)single.in + ((int64_t)c << single.nbits);
            S.out = (T *)single.out + ((int64_t)c << single.nbits);
        }
        // states have at least one qubit: tiles hold whole pairs
        const int tb = S.nbits < TBITS ? S.nbits : TBITS;
        const int tsize = 1 << tb;
        const uint32_t base = (uint32_t)w.tile << tb;
        const T *__restrict__ in = (const T *)S.in;
        T *__restrict__ out = (T *)S.out;
        const bool add_out = S.prog_hi_off >= 0;
        const bool bz = beta_zero && !add_out;
        const T beff = add_out ? mk(1.0f, 0.0f) : beta;
        T xv[EPT];
#pragma unroll
        for (int j = 0; j < EPT / 2; ++j) {
            const int e = 2 * (tid + j * NTH);
            if (e < tsize) {
                qp_load<true>((const char *)(in + base + e), xv[2 * j], xv[2 * j + 1]);
            } else {
                xv[2 * j] = czero<T>();
                xv[2 * j + 1] = czero<T>();
            }
        }
        T *tile = (T *)lds_raw + (par << TBITS);
        par ^= (int)dbuf;
        if (!dbuf) __syncthreads();  // the previous item's tile reads are finished
#pragma unroll
        for (int j = 0; j < EPT / 2; ++j) {
            qb_f4 v;
            v.x = xv[2 * j].x; v.y = xv[2 * j].y; v.z = xv[2 * j + 1].x; v.w = xv[2 * j + 1].y;
            *(qb_f4 *)(tile + 2 * (tid + j * NTH)) = v;
        }
        __syncthreads();
        T acc[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) acc[k] = czero<T>();
        qb_run_pair<NTH, EPT>(prog + S.prog_off, vals, tile, in, base, (uint32_t)tsize - 1u, tb, xv, acc);
#pragma unroll
        for (int j = 0; j < EPT / 2; ++j) {
            const int e = 2 * (tid + j * NTH);
            if (e < tsize) {
                T rr[2];
                T old[2];
                if (!bz) qp_load<true>((const char *)(out + base + e), old[0], old[1]);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    T v = cmul(alpha, acc[2 * j + h]);
                    if (!gamma_zero) v = cfma(gamma, xv[2 * j + h], v);
                    if (!bz) v = cfma(beff, old[h], v);
                    rr[h] = v;
                }
                qb_f4 o;
                o.x = rr[0].x; o.y = rr[0].y; o.z = rr[1].x; o.w = rr[1].y;
                __builtin_nontemporal_store(o, (QOB_GLOBAL qb_f4 *)(out + base + e));
            }
        }
    }
}

// Second pass: flips of bits above the tile.  With h = nbits - TBITS high
// bits and c = TBITS - h low bits, tile m holds the 2^TBITS basis states
//   J(q) = (q >> c) << TBITS | m << c | (q & (2^c - 1)),
// i.e. every combination of the high bits times a contiguous run of 2^c
// states, so a flip of high bits (and of bits below c) stays inside the
// tile.  Writes out = alpha * sum_high + beta * out.
template <typename T, int NTH, int TBITS>
__global__ void __launch_bounds__(NTH) k_qubit_hi(const QubitStateDev *__restrict__ states,
                                                  const QubitItemDev *__restrict__ items, int nitems,
                                                  const int32_t *__restrict__ prog, const T *__restrict__ vals,
                                                  double2 alpha_, double2 beta_, int flags, QubitStateDev single,
                                                  int ntiles1, QubitSplit xs) {
    constexpr int EPT = (1 << TBITS) / NTH;
    extern __shared__ unsigned char lds_raw[];
    T *tile = (T *)lds_raw;
    const T alpha = to_t<T>(alpha_), beta = to_t<T>(beta_);
    const bool beta_zero = flags & 1;
    const int tid = threadIdx.x;
    const ItemWalk wk = qb_walk(flags, xs, nitems);
    for (int it = wk.it; it < wk.end; it += wk.step) {
        QubitItemDev w;
        QubitStateDev S;
        if (states) {
            w = items[it];
            S = states[w.state];
        } else {
            const int cidx = it / ntiles1;
            w.state = cidx;
            w.tile = it - cidx * ntiles1;
            S = single;
            S.in = (const T *)single.in + ((int64_t)cidx << single.nbits);
            S.out = (T *)single.out + ((int64_t)cidx << single.nbits);
        }
        const int h = S.nbits - TBITS;
        const int c = TBITS - h;
        const HiMap<TBITS> m{(uint32_t)w.tile << c, (1u << c) - 1u, c};
        const T *__restrict__ in = (const T *)S.in;
        T *__restrict__ out = (T *)S.out;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const uint32_t q = (uint32_t)(tid + k * NTH);
            tile[q] = ldg(in + m.J(q));
        }
        __syncthreads();
        T acc[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) acc[k] = czero<T>();
        const T(&xv)[EPT] = acc;  // not read: the second pass has no register flips
        qb_run<T, NTH, EPT, false>(prog + S.prog_hi_off, vals, tile, in, m, xv, acc);
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const uint32_t J = m.J((uint32_t)(tid + k * NTH));
            T r = cmul(alpha, acc[k]);
            if (!beta_zero) r = cfma(beta, ldg(out + J), r);
            stg(out + J, r);
        }
    }
}

// sum_J conj(u[J]) (H v)[J] for one state (expect / variance): the same tile
// walk as k_qubit with the store replaced by an FP64 reduction; one partial
// per workgroup (fixed item order, so the result is deterministic).  High
// flips read partner tiles from v in global memory.
template <typename T, int NTH, int TBITS>
__global__ void __launch_bounds__(NTH) k_qubit_dot(QubitStateDev S, const T *__restrict__ u, int ntiles,
                                                   const int32_t *__restrict__ prog, const T *__restrict__ vals,
                                                   double2 *__restrict__ partials) {
    constexpr int EPT = (1 << TBITS) / NTH;
    extern __shared__ unsigned char lds_raw[];
    T *tile = (T *)lds_raw;
    __shared__ double2 red[NTH / 64];
    const int tid = threadIdx.x;
    const int tb = S.nbits < TBITS ? S.nbits : TBITS;
    const int tsize = 1 << tb;
    const T *__restrict__ v = (const T *)S.in;
    double rx = 0, ry = 0;
    for (int it = blockIdx.x; it < ntiles; it += gridDim.x) {
        const uint32_t base = (uint32_t)it << tb;
        T xv[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = tid + k * NTH;
            xv[k] = e < tsize ? ldg(v + base + e) : czero<T>();
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < EPT; ++k) tile[tid + k * NTH] = xv[k];
        __syncthreads();
        T acc[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) acc[k] = czero<T>();
        const LoMap m{base, (uint32_t)tsize - 1u, tb};
        qb_run<T, NTH, EPT, true>(prog + S.prog_off, vals, tile, v, m, xv, acc);
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = tid + k * NTH;
            if (e < tsize) {
                const T a = u[base + e];
                rx += (double)a.x * acc[k].x + (double)a.y * acc[k].y;
                ry += (double)a.x * acc[k].y - (double)a.y * acc[k].x;
            }
        }
    }
    // wave sums, then one lane adds the waves in a fixed order
    for (int off = 32; off > 0; off >>= 1) {
        rx += __shfl_down(rx, off, 64);
        ry += __shfl_down(ry, off, 64);
    }
    if ((tid & 63) == 0) red[tid >> 6] = make_double2(rx, ry);
    __syncthreads();
    if (tid == 0) {
        double2 t = make_double2(0, 0);
        for (int w = 0; w < NTH / 64; ++w) {
            t.x += red[w].x;
            t.y += red[w].y;
        }
        partials[blockIdx.x] = t;
    }
}

// Elementwise batch update out += a * in over the same items (the axpy of a
// time step for every state of a batch).
template <typename T, int NTH>
__global__ void __launch_bounds__(NTH) k_qubit_axpy(const QubitStateDev *__restrict__ states,
                                                    const QubitItemDev *__restrict__ items, int nitems, int tbits,
                                                    double2 a_) {
    const T a = to_t<T>(a_);
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const QubitItemDev w = items[it];
        const QubitStateDev S = states[w.state];
        const int tb = S.nbits < tbits ? S.nbits : tbits;
        const uint32_t base = (uint32_t)w.tile << tb;
        const T *__restrict__ in = (const T *)S.in;
        T *__restrict__ out = (T *)S.out;
        for (int e = threadIdx.x; e < (1 << tb); e += NTH)
            stg(out + base + e, cfma(a, ldg(in + base + e), ldg(out + base + e)));
    }
}

}  // namespace

// threads per workgroup; 0: by element type and tile, 4 elements per lane in
// ComplexF64 (1024 threads for its 2^12-element tiles) and 8 in ComplexF32
// (256 threads for 2^11), measured fastest on the 256-chain batch
int g_qubit_nth = 0;
static int qubit_nth(qob_dtype dt, int tbits) {
    if (g_qubit_nth) return g_qubit_nth;
    return std::min(1024, std::max(256, (1 << tbits) >> (dt == QOB_C64 ? 2 : 3)));
}
// 0: workgroup b walks items b, b + G, ...; 1: the same rounds of G items with
// each XCD given a contiguous run of every round; 2: each XCD walks its own
// item range for the whole launch (a batch's states are assigned to XCDs
// whole; the columns form splits its items into 8 contiguous ranges).
int g_qubit_remap = 2;
// k >= 1: partner tiles of flips whose tile distance is at least the number
// of workgroups walking the XCD's range, halved k - 1 times, are read with
// non-temporal loads; 0: none
int g_qubit_ntfar = 1;
// 1: two tile buffers in LDS when they fit (one barrier per item instead of two)
int g_qubit_dbuf = 1;
// 1: ComplexF32 lanes hold pairs of adjacent elements (16-byte accesses;
// 12% faster on the 256-chain batch, the same bits)
int g_qubit_pair = 1;

// flags bits 4 / 8 and the item ranges of a launch with g workgroups
static int qubit_walk_flags(int nitems, unsigned g, const QubitSplit *xs, QubitSplit &use) {
    if (g_qubit_remap == 2 && g >= 8) {
        if (xs) {
            use = *xs;
        } else {
            for (int x = 0; x <= 8; ++x) use.off[x] = (int32_t)((int64_t)nitems * x / 8);
        }
        return 8;
    }
    for (int x = 0; x <= 8; ++x) use.off[x] = 0;
    return g_qubit_remap ? 4 : 0;
}
// 0: one workgroup per item when the items fit in 5 workgroups per CU (many small states: no
// persistent loop, every item starts at once); otherwise a persistent grid of 1024 / nth
// workgroups per CU (16 waves), the most that fit by registers (84 VGPRs in ComplexF64, 54 in
// ComplexF32 at 4 and 8 elements per lane). Each wave keeps two flip groups' partner loads in
// flight; the tiles in flight per XCD (2 MB in ComplexF64: 32 of 2^12 elements) stay well
// inside its 4 MB L2, which must also hold the partner tiles read again.
int g_qubit_grid = 0;

static unsigned qubit_grid(int nth, int nitems) {
    const int ncu = device_cu_count();
    const int64_t g = g_qubit_grid > 0           ? g_qubit_grid
                      : nitems <= (int64_t)ncu * 5 ? nitems
                                                   : (int64_t)ncu * std::max(1, 1024 / nth);
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, nitems));
}

// log2 of the tile; 0: 12 for ComplexF64 (one partner-tile read per element
// fewer than 2^11: 3.1 reads per element instead of 4.1 on the 256-chain
// batch, 10% faster), 11 for ComplexF32 (bound by issue, not traffic: larger
// tiles measured slower)
int g_qubit_tbits = 0;
int qubit_tbits(qob_dtype dt) { return g_qubit_tbits ? g_qubit_tbits : (dt == QOB_C64 ? 12 : 11); }

static hipError_t launch_qubit_any(qob_dtype dt, const QubitStateDev *states, const QubitItemDev *items, int nitems,
                                   int tbits, const int32_t *prog, const void *vals, double2 alpha, double2 beta,
                                   double2 gamma, hipStream_t s, QubitStateDev single, int ntiles1,
                                   const QubitSplit *xs);

hipError_t launch_qubit(qob_dtype dt, const QubitStateDev *states, const QubitItemDev *items, int nitems, int tbits,
                        const int32_t *prog, const void *vals, double2 alpha, double2 beta, double2 gamma,
                        hipStream_t s, const QubitSplit *xs) {
    return launch_qubit_any(dt, states, items, nitems, tbits, prog, vals, alpha, beta, gamma, s, QubitStateDev(), 1,
                            xs);
}

hipError_t launch_qubit_cols(qob_dtype dt, QubitStateDev single, int64_t ncols, int tbits, const int32_t *prog,
                             const void *vals, double2 alpha, double2 beta, double2 gamma, hipStream_t s) {
    const int tb = std::min(tbits, (int)single.nbits);
    const int64_t ntiles = (int64_t)1 << (single.nbits - tb);
    const int64_t nitems = ntiles * ncols;
    if (nitems >= (1ll << 31)) return hipErrorInvalidValue;
    return launch_qubit_any(dt, nullptr, nullptr, (int)nitems, tbits, prog, vals, alpha, beta, gamma, s, single,
                            (int)ntiles, nullptr);
}

static hipError_t launch_qubit_any(qob_dtype dt, const QubitStateDev *states, const QubitItemDev *items, int nitems,
                                   int tbits, const int32_t *prog, const void *vals, double2 alpha, double2 beta,
                                   double2 gamma, hipStream_t s, QubitStateDev single, int ntiles1,
                                   const QubitSplit *xs) {
    if (nitems == 0) return hipSuccess;
    if (tbits < 10) tbits = 10;
    if (tbits > 13) tbits = 13;
    const size_t tile_bytes = (size_t)(1 << tbits) * (dt == QOB_C64 ? 16 : 8);
    if (tile_bytes > 160 * 1024) return hipErrorInvalidValue;
    const int nth = qubit_nth(dt, tbits);
    const unsigned g = qubit_grid(nth, nitems);
    // a second buffer only pays where a workgroup walks several items
    const bool dbuf = g_qubit_dbuf && 2 * tile_bytes <= 160 * 1024 && (int64_t)g < nitems;
    const size_t lds = dbuf ? 2 * tile_bytes : tile_bytes;
    QubitSplit ws;
    const int flags = (beta.x == 0.0 && beta.y == 0.0 ? 1 : 0) | (gamma.x == 0.0 && gamma.y == 0.0 ? 2 : 0) |
                      qubit_walk_flags(nitems, g, xs, ws) | (g_qubit_ntfar ? 16 | ((g_qubit_ntfar - 1) << 5) : 0) |
                      (dbuf ? 256 : 0);
#define QOB_QB(TT, NT, TB)                                                                                    \
    do {                                                                                                      \
        auto kfn = k_qubit<TT, NT, TB>;                                                                       \
        static size_t lds_set = 0; /* raise the limit once, outside any stream capture */                     \
        if (lds > lds_set) {                                                                                  \
            hipError_t e_ = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                (int)(160 * 1024));                                           \
            if (e_ != hipSuccess) return e_;                                                                  \
            lds_set = 160 * 1024;                                                                             \
        }                                                                                                     \
        hipLaunchKernelGGL(kfn, dim3(g), dim3(NT), lds, s, states, items, nitems, prog, (const TT *)vals,     \
                           alpha, beta, gamma, flags, single, ntiles1, ws);                                  \
        return hipGetLastError();                                                                             \
    } while (0)
#define QOB_QB_NT(TT, TB)                                  \
    do {                                                   \
        if (nth == 512) QOB_QB(TT, 512, TB);                 \
        else if (nth == 1024) QOB_QB(TT, 1024, TB);          \
        else QOB_QB(TT, 256, TB);                          \
    } while (0)
#define QOB_QB_TB(TT)                          \
    do {                                       \
        if (tbits <= 10) QOB_QB_NT(TT, 10);    \
        else if (tbits == 11) QOB_QB_NT(TT, 11); \
        else if (tbits == 12) QOB_QB_NT(TT, 12); \
        else QOB_QB_NT(TT, 13);                \
    } while (0)
    if (dt == QOB_C64) QOB_QB_TB(double2);
    if (g_qubit_pair && (1 << tbits) / nth >= 2) {
#define QOB_QP(NT, TB)                                                                                        \
    do {                                                                                                      \
        auto kfn = k_qubit_pair<NT, TB>;                                                                      \
        static size_t lds_set = 0;                                                                            \
        if (lds > lds_set) {                                                                                  \
            hipError_t e_ = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                (int)(160 * 1024));                                           \
            if (e_ != hipSuccess) return e_;                                                                  \
            lds_set = 160 * 1024;                                                                             \
        }                                                                                                     \
        hipLaunchKernelGGL(kfn, dim3(g), dim3(NT), lds, s, states, items, nitems, prog, (const float2 *)vals, \
                           alpha, beta, gamma, flags, single, ntiles1, ws);                                   \
        return hipGetLastError();                                                                             \
    } while (0)
        // (at least two elements per lane, checked above)
        if (tbits <= 10) {
            if (nth == 256) QOB_QP(256, 10);
            if (nth == 512) QOB_QP(512, 10);
        } else if (tbits == 11) {
            if (nth == 256) QOB_QP(256, 11);
            if (nth == 512) QOB_QP(512, 11);
            if (nth == 1024) QOB_QP(1024, 11);
        } else if (tbits == 12) {
            if (nth == 256) QOB_QP(256, 12);
            if (nth == 512) QOB_QP(512, 12);
            if (nth == 1024) QOB_QP(1024, 12);
        } else {
            if (nth == 256) QOB_QP(256, 13);
            if (nth == 512) QOB_QP(512, 13);
            if (nth == 1024) QOB_QP(1024, 13);
        }
#undef QOB_QP
    }
    QOB_QB_TB(float2);
#undef QOB_QB_TB
#undef QOB_QB_NT
#undef QOB_QB
    return hipGetLastError();
}

static hipError_t launch_qubit_hi_any(qob_dtype dt, const QubitStateDev *states, const QubitItemDev *items,
                                      int nitems, int tbits, const int32_t *prog, const void *vals, double2 alpha,
                                      double2 beta, hipStream_t s, QubitStateDev single, int ntiles1,
                                      const QubitSplit *xs) {
    if (nitems == 0) return hipSuccess;
    if (tbits < 10 || tbits > 13) return hipErrorInvalidValue;
    const size_t lds = (size_t)(1 << tbits) * (dt == QOB_C64 ? 16 : 8);
    const int nth = qubit_nth(dt, tbits);
    const unsigned g = qubit_grid(nth, nitems);
    QubitSplit ws;
    const int flags = (beta.x == 0.0 && beta.y == 0.0 ? 1 : 0) | qubit_walk_flags(nitems, g, xs, ws);
#define QOB_QH(TT, NT, TB)                                                                                    \
    do {                                                                                                      \
        auto kfn = k_qubit_hi<TT, NT, TB>;                                                                    \
        static size_t lds_set = 0;                                                                            \
        if (lds > lds_set) {                                                                                  \
            hipError_t e_ = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                (int)(160 * 1024));                                           \
            if (e_ != hipSuccess) return e_;                                                                  \
            lds_set = 160 * 1024;                                                                             \
        }                                                                                                     \
        hipLaunchKernelGGL(kfn, dim3(g), dim3(NT), lds, s, states, items, nitems, prog, (const TT *)vals,     \
                           alpha, beta, flags, single, ntiles1, ws);                                          \
        return hipGetLastError();                                                                             \
    } while (0)
#define QOB_QH_NT(TT, TB)                                  \
    do {                                                   \
        if (nth == 512) QOB_QH(TT, 512, TB);                 \
        else if (nth == 1024) QOB_QH(TT, 1024, TB);          \
        else QOB_QH(TT, 256, TB);                          \
    } while (0)
#define QOB_QH_TB(TT)                            \
    do {                                         \
        if (tbits == 10) QOB_QH_NT(TT, 10);      \
        else if (tbits == 11) QOB_QH_NT(TT, 11); \
        else if (tbits == 12) QOB_QH_NT(TT, 12); \
        else QOB_QH_NT(TT, 13);                  \
    } while (0)
    if (dt == QOB_C64) QOB_QH_TB(double2);
    else QOB_QH_TB(float2);
#undef QOB_QH_TB
#undef QOB_QH_NT
#undef QOB_QH
}

hipError_t launch_qubit_hi(qob_dtype dt, const QubitStateDev *states, const QubitItemDev *items, int nitems, int tbits,
                           const int32_t *prog, const void *vals, double2 alpha, double2 beta, hipStream_t s,
                           const QubitSplit *xs) {
    return launch_qubit_hi_any(dt, states, items, nitems, tbits, prog, vals, alpha, beta, s, QubitStateDev(), 1, xs);
}

hipError_t launch_qubit_hi_cols(qob_dtype dt, QubitStateDev single, int64_t ncols, int tbits, const int32_t *prog,
                                const void *vals, double2 alpha, double2 beta, hipStream_t s) {
    if (single.nbits <= tbits || single.prog_hi_off < 0) return hipSuccess;
    const int64_t ntiles = (int64_t)1 << (single.nbits - tbits);
    const int64_t nitems = ntiles * ncols;
    if (nitems >= (1ll << 31)) return hipErrorInvalidValue;
    return launch_qubit_hi_any(dt, nullptr, nullptr, (int)nitems, tbits, prog, vals, alpha, beta, s, single,
                               (int)ntiles, nullptr);
}

int qubit_dot_blocks(int64_t ntiles) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(ntiles, (int64_t)device_cu_count() * 2));
}

hipError_t launch_qubit_dot(qob_dtype dt, QubitStateDev single, const void *u, int tbits, const int32_t *prog,
                            const void *vals, double2 *partials, hipStream_t s) {
    if (tbits < 10 || tbits > 13) return hipErrorInvalidValue;
    const int tb = std::min(tbits, (int)single.nbits);
    const int64_t ntiles = (int64_t)1 << (single.nbits - tb);
    if (ntiles >= (1ll << 31)) return hipErrorInvalidValue;
    const int g = qubit_dot_blocks(ntiles);
    const size_t lds = (size_t)(1 << tbits) * (dt == QOB_C64 ? 16 : 8);
#define QOB_QD(TT, TB)                                                                                          \
    do {                                                                                                        \
        constexpr int NTD = TB >= 12 ? 1024 : 512; /* 4 elements per lane up to 2^12-element tiles */          \
        auto kfn = k_qubit_dot<TT, NTD, TB>;                                                                    \
        static std::atomic<int> lds_set{0};                                                                     \
        if (!lds_set.load()) {                                                                                  \
            hipError_t e_ = hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize,  \
                                                150 * 1024);                                                    \
            if (e_ != hipSuccess) return e_;                                                                    \
            lds_set.store(1);                                                                                   \
        }                                                                                                       \
        hipLaunchKernelGGL(kfn, dim3(g), dim3(NTD), lds, s, single, (const TT *)u, (int)ntiles, prog,           \
                           (const TT *)vals, partials);                                                         \
        return hipGetLastError();                                                                               \
    } while (0)
    if (dt == QOB_C64) {
        if (tbits == 10) QOB_QD(double2, 10);
        else if (tbits == 11) QOB_QD(double2, 11);
        else if (tbits == 12) QOB_QD(double2, 12);
        else QOB_QD(double2, 13);
    } else {
        if (tbits == 10) QOB_QD(float2, 10);
        else if (tbits == 11) QOB_QD(float2, 11);
        else if (tbits == 12) QOB_QD(float2, 12);
        else QOB_QD(float2, 13);
    }
#undef QOB_QD
}

hipError_t launch_qubit_axpy(qob_dtype dt, const QubitStateDev *states, const QubitItemDev *items, int nitems,
                             int tbits, double2 a, hipStream_t s) {
    if (nitems == 0) return hipSuccess;
    const unsigned g = qubit_grid(256, nitems);
    if (dt == QOB_C64)
        hipLaunchKernelGGL((k_qubit_axpy<double2, 256>), dim3(g), dim3(256), 0, s, states, items, nitems, tbits, a);
    else
        hipLaunchKernelGGL((k_qubit_axpy<float2, 256>), dim3(g), dim3(256), 0, s, states, items, nitems, tbits, a);
    return hipGetLastError();
}

}  // namespace qob
