// Host side of libqob_hip.so: contexts, operator handles, planning and the
// mul! dispatch table of QuantumOpticsBase.jl, re-expressed as launches of
// the kernels in k_*.hip.
//
// Reference dispatch (Julia multiple dispatch on 5-argument mul!):
//   Operator x Operator/Ket/Bra      src/operators_dense.jl:394-396
//   SparseOperator x ...             src/operators_sparse.jl:199-202
//   LazyTensor x Ket/Bra/Operator    src/operators_lazytensor.jl:539-609
//   LazySum x Ket/Bra/Operator       src/operators_lazysum.jl:189-238
// Planning that the reference does per call (identity filtering, isometries,
// strides, temporaries) is done once per operator and cached on its handle.
#include "qob_internal.h"

#include <algorithm>
#include <array>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

using cd = std::complex<double>;

namespace qob {

extern int g_sell_variant, g_sell_cj, g_sell_jchunk, g_csc_rb, g_csc_cc, g_csc_remap, g_nt_store, g_copy_variant;
int g_fused_lt = 1;  // fused MFMA column kernel for dense LazyTensors
extern int g_qubit_nth, g_qubit_grid, g_qubit_remap, g_qubit_ntfar, g_qubit_dbuf, g_qubit_pair;
// order of a run's tiles in a batch (batches created after it is set):
// 0 index order; 1 binary-reflected Gray code of the index, so that tiles
// walked in turn differ in one bit and the partner runs of the high bits
// follow each other; 2 the same on the index bits above the low 6 only
int g_qubit_order = 1;
extern int g_mode_kernel, g_fused_gpw, g_csc_rows, g_sell_unr, g_csc_order, g_fused_tiles, g_fused_resident;
extern int g_fused_reg;
extern int g_sell_align, g_csc_waves, g_csc_sweep, g_sell_cjchunk, g_csc_gkernel, g_csc_gremap, g_sell_cremap;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

static qob_status vfail(qob_status st, const char *fmt, va_list ap) {
    char buf[512];
    vsnprintf(buf, sizeof buf, fmt, ap);
    g_err = buf;
    return st;
}

static qob_status fail(qob_status st, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfail(st, fmt, ap);
    va_end(ap);
    return st;
}

qob_status set_error(qob_status st, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfail(st, fmt, ap);
    va_end(ap);
    return st;
}

#define QOB_HIP(x)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) return fail(QOB_EHIP, "%s: %s", #x, hipGetErrorString(e_));       \
    } while (0)
#define QOB_TRY(x)                          \
    do {                                    \
        qob_status st_ = (x);               \
        if (st_ != QOB_OK) return st_;      \
    } while (0)

static inline double2 d2(cd z) { return make_double2(z.real(), z.imag()); }
static inline double2 d2(qob_c64 z) { return make_double2(z.re, z.im); }
static inline cd cdv(qob_c64 z) { return cd(z.re, z.im); }
static inline bool is_zero(cd z) { return z.real() == 0.0 && z.imag() == 0.0; }
static inline size_t esize(qob_dtype dt) { return dt == QOB_C64 ? 16 : 8; }

// -------------------------------------------------------------- device buf
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    qob_status alloc(size_t b) {
        if (b <= bytes && p) return QOB_OK;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            bytes = 0;
        }
        if (b == 0) b = 16;
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            return fail(QOB_ENOMEM, "hipMalloc of %zu bytes failed", b);
        }
        bytes = b;
        return QOB_OK;
    }
    qob_status upload(const void *src, size_t b) {
        QOB_TRY(alloc(b));
        if (b) QOB_HIP(hipMemcpy(p, src, b, hipMemcpyHostToDevice));
        return QOB_OK;
    }
    // hand the allocation over (freed when the last owner lets go); empty after
    std::shared_ptr<void> release() {
        std::shared_ptr<void> r;
        if (p) r = std::shared_ptr<void>(p, [](void *q) { (void)hipFree(q); });
        p = nullptr;
        bytes = 0;
        return r;
    }
};

// Upload complex values in the requested dtype.
static qob_status upload_vals(DevBuf &buf, const std::vector<cd> &v, qob_dtype dt) {
    if (dt == QOB_C64) {
        std::vector<double2> h(v.size());
        for (size_t i = 0; i < v.size(); ++i) h[i] = d2(v[i]);
        return buf.upload(h.data(), h.size() * sizeof(double2));
    }
    std::vector<float2> h(v.size());
    for (size_t i = 0; i < v.size(); ++i) h[i] = make_float2((float)v[i].real(), (float)v[i].imag());
    return buf.upload(h.data(), h.size() * sizeof(float2));
}

// ------------------------------------------------------------- row lists
// Host sparse matrix as lists of rows: row i holds (idx[p], val[p]) for
// p in [ptr[i], ptr[i+1]).
struct RowLists {
    int64_t nr = 0, nc = 0;
    std::vector<int32_t> ptr, idx;
    std::vector<cd> val;
    bool monomial() const {
        for (int64_t i = 0; i < nr; ++i)
            if (ptr[i + 1] - ptr[i] > 1) return false;
        return true;
    }
};

// Transpose of row lists (rows of X -> rows of X^T), optionally conjugated.
static RowLists transpose_rows(const RowLists &X, bool conj) {
    RowLists T;
    T.nr = X.nc;
    T.nc = X.nr;
    T.ptr.assign(T.nr + 1, 0);
    for (int32_t c : X.idx) T.ptr[c + 1]++;
    for (int64_t i = 0; i < T.nr; ++i) T.ptr[i + 1] += T.ptr[i];
    T.idx.resize(X.idx.size());
    T.val.resize(X.val.size());
    std::vector<int32_t> pos(T.ptr.begin(), T.ptr.end() - 1);
    for (int64_t r = 0; r < X.nr; ++r) {
        for (int32_t p = X.ptr[r]; p < X.ptr[r + 1]; ++p) {
            const int32_t c = X.idx[p];
            T.idx[pos[c]] = (int32_t)r;
            T.val[pos[c]] = conj ? std::conj(X.val[p]) : X.val[p];
            pos[c]++;
        }
    }
    return T;
}

static RowLists identity_rows(int64_t m, int64_t n) {
    RowLists R;
    R.nr = m;
    R.nc = n;
    R.ptr.assign(m + 1, 0);
    const int64_t d = std::min(m, n);
    for (int64_t i = 0; i < m; ++i) {
        R.ptr[i + 1] = R.ptr[i] + (i < d ? 1 : 0);
        if (i < d) {
            R.idx.push_back((int32_t)i);
            R.val.push_back(cd(1, 0));
        }
    }
    return R;
}

// Variant of an operator's matrix: bit0 = transpose, bit1 = conjugate.
// N = 0, T = 1, conj = 2, adjoint H = 3; composition is XOR.
enum { V_N = 0, V_T = 1, V_CJ = 2, V_H = 3 };

// --------------------------------------------------------------- sparse op
struct RefreshMap;
// src: for every value slot of the table, the entry of the row lists it was
// built from (-1: a padding zero); rmap: the slot values as linear functions
// of a LazySum's coefficients (materialised sums under device-held
// coefficients, built on first need)
struct SellTab {
    DevBuf off, col, val, rows;
    SellDev dev;
    std::vector<int64_t> src;
    std::shared_ptr<RefreshMap> rmap;
};
struct CscTab {
    DevBuf ptr, row, val, ord, gint, gcoef;
    CscDev dev;
    std::vector<int64_t> gsrc;  // of gcoef (val is the row lists' values in order)
    std::shared_ptr<RefreshMap> rmap, grmap;
};

// A materialised LazySum's entries as linear functions of its coefficients:
// entry e of the summed CSR matrix is sum over c in [cptr[e], cptr[e + 1]) of
// coef[cterm[c]] * cbase[c]; csc2csr maps the CSC entries (transpose) to
// CSR entries.  dcoef / dcoef_on are the owner's device coefficients.
struct MatContrib {
    std::vector<int64_t> cptr;
    std::vector<int32_t> cterm;
    std::vector<std::complex<double>> cbase;
    std::vector<int64_t> csc2csr;
    const void *const *dcoef = nullptr;
    const bool *dcoef_on = nullptr;
};

struct SparseData {
    int64_t m = 0, n = 0;
    RowLists csc;  // columns of M as lists (rows of M^T)
    RowLists csr;  // rows of M
    std::mutex mu;
    std::map<int, std::shared_ptr<SellTab>> sell;   // key: var*2 + dt
    std::map<int, std::shared_ptr<CscTab>> cols;    // key: var*2 + dt
    std::shared_ptr<MatContrib> mc;                 // materialised LazySum only

    // rows of op_v(M)
    RowLists rows(int v) const {
        const bool t = v & 1, c = (v >> 1) & 1;
        const RowLists &src = t ? csc : csr;
        if (!c) return src;
        RowLists r = src;
        for (auto &z : r.val) z = std::conj(z);
        return r;
    }
};

// Slots of a 64-row slice.  Packed: a row's entries fill slots 0, 1, ... in
// column order, so one slot mixes diagonals (offsets column - row) across
// the rows of the slice and one load instruction of the kernel touches
// several separate runs of B.  Diagonal-aligned (when the slice holds at
// most 16 distinct offsets and no more than twice its widest row): slot k
// holds offset d_k in every row that has it, so the lanes of a slot read one
// contiguous run of a column of B.  Rows without an entry on d_k keep the
// empty marker (-1) and read nothing.
int g_sell_align = 1;

// Closed blocks.  For a square X whose pattern (rows i and k joined when
// X[i,k] != 0, either way round) splits into connected components of at most
// 256 rows, whole components are packed into blocks of 256 rows (components
// in order of their first row, a new block when the next one does not fit;
// rows ascending within a block; a block is 4 slices of 64 lanes).  Every
// row of B that a block reads is then the row one of its own lanes owns:
// each lane loads its own row of B once per column into the workgroup's LDS
// and the entries read from there, so B is read exactly once.  Such
// components are the blocks of a conserved quantity (the chains of a
// Jaynes-Cummings-like exchange term, for example).  Returns false
// (contiguous slices) for any other X.
//
// Line-aligned form (ComplexF64, sell_cline = 1): rows are also joined
// within aligned groups of 8 (the 128-byte line of a column of R), so every
// run of rows in a block is whole lines and no two blocks write into one
// line.  The components then grow; they are packed into blocks of 320 rows
// (5 slices), used when that fills the blocks at least 90%.
//
// Row-pair form (ComplexF32, sell_cpair = 1): rows 2i and 2i + 1 are joined,
// so every block holds whole pairs at even lanes and a thread can move two
// rows with one 16-byte access per column.  Blocks of 256 rows, else 320,
// whichever is first filled at least 90%; else the plain 256-row form.
int g_sell_closed = 1;
int g_sell_cline = 1;
int g_sell_cpair = 1;

// lanes (block-major, BLK per block, -1 idle) of the closed layout with rows
// joined in aligned groups of grp; false if a component exceeds BLK rows
static bool closed_lanes(const RowLists &X, int grp, int BLK, std::vector<int32_t> &lanes) {
    const int64_t m = X.nr;
    std::vector<int32_t> parent((size_t)m);
    for (int64_t i = 0; i < m; ++i) parent[i] = (int32_t)i;
    auto find = [&](int32_t a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    auto unite = [&](int32_t a, int32_t b) {
        a = find(a);
        b = find(b);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    };
    for (int64_t i = 0; i < m; ++i)
        for (int32_t q = X.ptr[i]; q < X.ptr[i + 1]; ++q) unite((int32_t)i, X.idx[q]);
    if (grp > 1)
        for (int64_t i = 0; i < m; ++i)
            if (i % grp) unite((int32_t)i, (int32_t)(i - i % grp));
    std::vector<int32_t> size((size_t)m, 0);
    for (int64_t i = 0; i < m; ++i) {
        if (++size[find((int32_t)i)] > BLK) return false;
    }
    // components in order of their first (smallest) row; rows of each component ascending
    std::vector<int32_t> start((size_t)m + 1, 0), cur, rows_of((size_t)m);
    for (int64_t i = 0; i < m; ++i) start[find((int32_t)i) + 1]++;
    for (int64_t i = 0; i < m; ++i) start[i + 1] += start[i];
    cur.assign(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < m; ++i) rows_of[cur[find((int32_t)i)]++] = (int32_t)i;
    lanes.clear();
    int used = BLK;
    for (int64_t r = 0; r < m; ++r) {
        if (start[r + 1] == start[r]) continue;  // r is not a component root
        const int sz = start[r + 1] - start[r];
        if (used + sz > BLK) {
            lanes.insert(lanes.end(), BLK, -1);
            used = 0;
        }
        const size_t base = lanes.size() - BLK;
        for (int k = 0; k < sz; ++k) lanes[base + used + k] = rows_of[start[r] + k];
        used += sz;
    }
    const int nblocks = (int)(lanes.size() / BLK);
    for (int k = 0; k < nblocks; ++k) {
        auto b = lanes.begin() + (int64_t)k * BLK, e = b + BLK;
        std::sort(b, std::find(b, e, -1));
    }
    return true;
}

static bool build_sell_closed(const RowLists &X, qob_dtype dt, SellTab &S) {
    const int64_t m = X.nr;
    if (m != X.nc || m < 2 || m >= (1ll << 31)) return false;
    std::vector<int32_t> lanes;
    int spb = 0;
    auto filled = [&](int grp, int blk) {
        return closed_lanes(X, grp, blk, lanes) && (double)m >= 0.9 * (double)lanes.size();
    };
    if (g_sell_cline && dt == QOB_C64 && filled(8, 320)) spb = 5;
    if (!spb && g_sell_cpair && dt == QOB_C32 && m % 2 == 0) {
        if (filled(2, 256)) spb = 4;
        else if (filled(2, 320)) spb = 5;
    }
    if (!spb) {
        if (!closed_lanes(X, 1, 256, lanes)) return false;
        spb = 4;
    }
    const int BLK = 64 * spb;
    const int nblocks = (int)(lanes.size() / BLK), nslices = nblocks * spb;
    std::vector<int32_t> lane_of((size_t)m, -1), off(nslices + 1, 0);
    for (size_t e = 0; e < lanes.size(); ++e)
        if (lanes[e] >= 0) lane_of[lanes[e]] = (int32_t)(e % BLK);
    for (int k = 0; k < nslices; ++k) {
        int w = 0;
        for (int l = 0; l < 64; ++l) {
            const int32_t r = lanes[(size_t)k * 64 + l];
            if (r >= 0) w = std::max(w, X.ptr[r + 1] - X.ptr[r]);
        }
        if (w > 16) return false;  // the kernel keeps a row's entries in registers
        off[k + 1] = off[k] + std::max(w, 1) * 64;
    }
    std::vector<int32_t> col(off[nslices], -1);
    std::vector<cd> val(off[nslices], cd(0, 0));
    S.src.assign(off[nslices], -1);
    for (int k = 0; k < nslices; ++k)
        for (int l = 0; l < 64; ++l) {
            const int32_t r = lanes[(size_t)k * 64 + l];
            if (r < 0) continue;
            int p = 0;
            for (int32_t q = X.ptr[r]; q < X.ptr[r + 1]; ++q, ++p) {
                const int64_t e = off[k] + (int64_t)p * 64 + l;
                col[e] = lane_of[X.idx[q]];
                val[e] = X.val[q];
                S.src[e] = q;
            }
        }
    if (S.off.upload(off.data(), off.size() * 4) != QOB_OK || S.col.upload(col.data(), col.size() * 4) != QOB_OK ||
        upload_vals(S.val, val, dt) != QOB_OK || S.rows.upload(lanes.data(), lanes.size() * 4) != QOB_OK)
        return false;
    S.dev.m = m;
    S.dev.n = X.nc;
    S.dev.nslices = nslices;
    S.dev.maxw = 0;
    for (int k = 0; k < nslices; ++k) S.dev.maxw = std::max(S.dev.maxw, (off[k + 1] - off[k]) / 64);
    S.dev.slice_off = (const int32_t *)S.off.p;
    S.dev.col = (const int32_t *)S.col.p;
    S.dev.val = S.val.p;
    S.dev.rows = (const int32_t *)S.rows.p;
    S.dev.spb = spb;
    S.dev.pairs = g_sell_cpair && dt == QOB_C32 && m % 2 == 0;
    for (size_t e = 0; e < lanes.size() && S.dev.pairs; e += 2)
        S.dev.pairs = lanes[e] < 0 ? lanes[e + 1] < 0 : lanes[e] % 2 == 0 && lanes[e + 1] == lanes[e] + 1;
    return true;
}

static qob_status build_sell(const RowLists &X, qob_dtype dt, SellTab &S) {
    if (g_sell_closed && build_sell_closed(X, dt, S)) return QOB_OK;
    const int64_t m = X.nr;
    const int nslices = (int)((m + 63) / 64);
    std::vector<int32_t> off(nslices + 1, 0);
    std::vector<std::vector<int64_t>> diag(nslices);  // aligned slices: their offsets, ascending
    for (int k = 0; k < nslices; ++k) {
        int w = 0;
        const int64_t r0 = (int64_t)k * 64, r1 = std::min<int64_t>(m, r0 + 64);
        for (int64_t r = r0; r < r1; ++r) w = std::max(w, X.ptr[r + 1] - X.ptr[r]);
        if (g_sell_align && w > 0) {
            std::vector<int64_t> d;
            for (int64_t r = r0; r < r1; ++r)
                for (int32_t q = X.ptr[r]; q < X.ptr[r + 1]; ++q) d.push_back((int64_t)X.idx[q] - r);
            std::sort(d.begin(), d.end());
            d.erase(std::unique(d.begin(), d.end()), d.end());
            if ((int)d.size() <= 16 && (int)d.size() <= 2 * w) {
                w = (int)d.size();
                diag[k] = std::move(d);
            }
        }
        off[k + 1] = off[k] + w * 64;
    }
    std::vector<int32_t> col(off[nslices], -1);
    std::vector<cd> val(off[nslices], cd(0, 0));
    S.src.assign(off[nslices], -1);
    for (int k = 0; k < nslices; ++k) {
        const std::vector<int64_t> &d = diag[k];
        for (int64_t r = (int64_t)k * 64; r < std::min<int64_t>(m, (int64_t)k * 64 + 64); ++r) {
            int p = 0;
            for (int32_t q = X.ptr[r]; q < X.ptr[r + 1]; ++q, ++p) {
                const int slot = d.empty() ? p
                                           : (int)(std::lower_bound(d.begin(), d.end(), (int64_t)X.idx[q] - r) -
                                                   d.begin());
                const int64_t e = off[k] + (int64_t)slot * 64 + (r - (int64_t)k * 64);
                col[e] = X.idx[q];
                val[e] = X.val[q];
                S.src[e] = q;
            }
        }
    }
    QOB_TRY(S.off.upload(off.data(), off.size() * 4));
    QOB_TRY(S.col.upload(col.data(), col.size() * 4));
    QOB_TRY(upload_vals(S.val, val, dt));
    S.dev.m = m;
    S.dev.n = X.nc;
    S.dev.nslices = nslices;
    S.dev.maxw = 0;
    for (int k = 0; k < nslices; ++k) S.dev.maxw = std::max(S.dev.maxw, (off[k + 1] - off[k]) / 64);
    S.dev.slice_off = (const int32_t *)S.off.p;
    S.dev.col = (const int32_t *)S.col.p;
    S.dev.val = S.val.p;
    return QOB_OK;
}

// Traversal order of the output columns for the right product R = B X.
// Output columns c and c' that share a row k of X both read column k of B;
// the pairs are the off-diagonal pattern of X^T X.  The kernel sweeps a row
// stripe of B per XCD across the columns in traversal order, so the second
// read of a column of B hits that XCD's L2 only if the two columns are close
// in the order.  Reverse Cuthill-McKee on the pattern of X^T X (one
// breadth-first pass per connected component, from a pseudo-peripheral
// column, neighbours by increasing degree) minimises that distance.  Returns
// an empty vector (natural order) when the natural order already keeps the
// distance within 64 columns, when the reordering does not at least halve it,
// or when X^T X would have more than 2^26 entries to visit.  *sweep_bw gets
// that distance for the order chosen (-1 when the graph was not built).
static std::vector<int32_t> csc_column_order(const RowLists &Xcols, int64_t xrows, int64_t *sweep_bw) {
    const int64_t n = Xcols.nr;
    *sweep_bw = 0;
    if (n < 2 || xrows < 1) return {};
    std::vector<int32_t> rp((size_t)xrows + 1, 0);  // rows of X as lists of columns
    for (int32_t r : Xcols.idx) rp[(size_t)r + 1]++;
    int64_t visits = 0;
    for (int64_t r = 0; r < xrows; ++r) {
        visits += (int64_t)rp[r + 1] * rp[r + 1];
        rp[r + 1] += rp[r];
    }
    if (visits > (1ll << 26)) {
        *sweep_bw = -1;
        return {};
    }
    std::vector<int32_t> rc(Xcols.idx.size()), cur(rp.begin(), rp.end() - 1);
    for (int64_t c = 0; c < n; ++c)
        for (int32_t p = Xcols.ptr[c]; p < Xcols.ptr[c + 1]; ++p) rc[cur[Xcols.idx[p]]++] = (int32_t)c;
    std::vector<int64_t> ap((size_t)n + 1, 0);
    std::vector<int32_t> adj, mark((size_t)n, -1);
    int64_t bw_nat = 0;
    for (int64_t c = 0; c < n; ++c) {
        for (int32_t p = Xcols.ptr[c]; p < Xcols.ptr[c + 1]; ++p) {
            const int32_t r = Xcols.idx[p];
            for (int32_t q = rp[r]; q < rp[r + 1]; ++q) {
                const int32_t c2 = rc[q];
                if (c2 != c && mark[c2] != (int32_t)c) {
                    mark[c2] = (int32_t)c;
                    adj.push_back(c2);
                    bw_nat = std::max<int64_t>(bw_nat, std::abs((int64_t)c2 - c));
                }
            }
        }
        ap[c + 1] = (int64_t)adj.size();
    }
    *sweep_bw = bw_nat;
    if (bw_nat <= 64) return {};
    auto deg = [&](int64_t c) { return ap[c + 1] - ap[c]; };
    // breadth-first levels from s inside its component: returns the last level
    std::vector<int32_t> lvl((size_t)n, -1), q;
    auto levels = [&](int32_t s, std::vector<int32_t> &comp) -> int {
        comp.clear();
        comp.push_back(s);
        lvl[s] = 0;
        int depth = 0;
        for (size_t h = 0; h < comp.size(); ++h) {
            const int32_t u = comp[h];
            depth = std::max(depth, lvl[u]);
            for (int64_t e = ap[u]; e < ap[u + 1]; ++e)
                if (lvl[adj[e]] < 0) {
                    lvl[adj[e]] = lvl[u] + 1;
                    comp.push_back(adj[e]);
                }
        }
        return depth;
    };
    std::vector<int32_t> order, comp, nb;
    order.reserve((size_t)n);
    std::vector<char> done((size_t)n, 0);
    for (int64_t c0 = 0; c0 < n; ++c0) {
        if (done[c0]) continue;
        // pseudo-peripheral start: move to a minimum-degree column of the last
        // level while that increases the depth (at most 4 moves)
        int32_t s = (int32_t)c0;
        int depth = levels(s, comp);
        for (int it = 0; it < 4; ++it) {
            int32_t best = -1;
            for (int32_t u : comp)
                if (lvl[u] == depth && (best < 0 || deg(u) < deg(best))) best = u;
            for (int32_t u : comp) lvl[u] = -1;
            const int d2 = levels(best, comp);
            if (d2 <= depth) {
                for (int32_t u : comp) lvl[u] = -1;
                levels(s, comp);
                break;
            }
            s = best;
            depth = d2;
        }
        for (int32_t u : comp) lvl[u] = -1;
        // Cuthill-McKee pass over the component
        const size_t first = order.size();
        order.push_back(s);
        done[s] = 1;
        for (size_t h = first; h < order.size(); ++h) {
            const int32_t u = order[h];
            nb.clear();
            for (int64_t e = ap[u]; e < ap[u + 1]; ++e)
                if (!done[adj[e]]) {
                    done[adj[e]] = 1;
                    nb.push_back(adj[e]);
                }
            std::sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) {
                return deg(a) != deg(b) ? deg(a) < deg(b) : a < b;
            });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    std::vector<int64_t> pos((size_t)n);
    for (int64_t k = 0; k < n; ++k) pos[order[k]] = k;
    int64_t bw = 0;
    for (int64_t c = 0; c < n; ++c)
        for (int64_t e = ap[c]; e < ap[c + 1]; ++e) bw = std::max<int64_t>(bw, std::abs(pos[c] - pos[adj[e]]));
    if (2 * bw > bw_nat) return {};
    *sweep_bw = bw;
    return order;
}

// Component groups for the right product R = B X.  For a square X whose
// pattern (output column c joined with every column k of B it reads) splits
// into connected components of at most 8 columns, each component becomes a
// group: its output columns (ascending), the columns of B they read
// (ascending), a mask of the stored entries and their values.  One wave per
// (group, 64-row block) loads each column of B the group reads once and
// forms every output column of the group from registers, so a column of B
// is read once per row block instead of once per output column that names
// it.  Entries of an output column are summed in ascending input column,
// the CSC order, so results equal the CSC kernel bit for bit.  Groups are
// listed by their first column.  Returns false for any other X.
int g_csc_groups = 1;

static bool build_csc_groups(const RowLists &Xcols, int64_t xrows, qob_dtype dt, CscTab &C) {
    const int64_t n = Xcols.nr;
    if (n != xrows || n < 2 || n >= (1ll << 31)) return false;
    std::vector<int32_t> parent((size_t)n);
    for (int64_t i = 0; i < n; ++i) parent[i] = (int32_t)i;
    auto find = [&](int32_t a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    for (int64_t c = 0; c < n; ++c)
        for (int32_t p = Xcols.ptr[c]; p < Xcols.ptr[c + 1]; ++p) {
            int32_t a = find((int32_t)c), b = find(Xcols.idx[p]);
            if (a != b) parent[std::max(a, b)] = std::min(a, b);
        }
    std::vector<int32_t> size((size_t)n, 0);
    int smax = 0;
    for (int64_t i = 0; i < n; ++i) smax = std::max(smax, ++size[find((int32_t)i)]);
    if (smax > 8) return false;
    const int GM = smax <= 2 ? 2 : smax <= 4 ? 4 : 8, NI = 4 + 2 * GM;
    std::vector<int32_t> start((size_t)n + 1, 0), cur, members((size_t)n);
    for (int64_t i = 0; i < n; ++i) start[find((int32_t)i) + 1]++;
    for (int64_t i = 0; i < n; ++i) start[i + 1] += start[i];
    cur.assign(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < n; ++i) members[cur[find((int32_t)i)]++] = (int32_t)i;
    std::vector<int32_t> gint;
    std::vector<cd> gcoef;
    std::vector<int64_t> gsrc;
    std::vector<int32_t> slot((size_t)n, -1);
    int ngroups = 0;
    for (int64_t r = 0; r < n; ++r) {
        const int sz = start[r + 1] - start[r];
        if (sz == 0) continue;
        const int32_t *mem = members.data() + start[r];
        // inputs: members some output of the group reads, ascending
        std::vector<int32_t> in;
        for (int k = 0; k < sz; ++k) slot[mem[k]] = -1;
        for (int k = 0; k < sz; ++k)
            for (int32_t p = Xcols.ptr[mem[k]]; p < Xcols.ptr[mem[k] + 1]; ++p) slot[Xcols.idx[p]] = 0;
        for (int k = 0; k < sz; ++k)
            if (slot[mem[k]] == 0) {
                slot[mem[k]] = (int32_t)in.size();
                in.push_back(mem[k]);
            }
        const size_t gi = gint.size();
        gint.resize(gi + NI, 0);
        gcoef.resize((size_t)(ngroups + 1) * GM * GM, cd(0, 0));
        gsrc.resize(gcoef.size(), -1);
        gint[gi] = sz;
        gint[gi + 1] = (int32_t)in.size();
        uint64_t mask = 0;
        for (int o = 0; o < sz; ++o) {
            gint[gi + 4 + o] = mem[o];
            for (int32_t p = Xcols.ptr[mem[o]]; p < Xcols.ptr[mem[o] + 1]; ++p) {
                const int q = slot[Xcols.idx[p]];
                const uint64_t bit = 1ull << (o * GM + q);
                if (mask & bit) return false;  // repeated entry: leave it to the CSC kernel
                mask |= bit;
                gcoef[(size_t)ngroups * GM * GM + (size_t)o * GM + q] = Xcols.val[p];
                gsrc[(size_t)ngroups * GM * GM + (size_t)o * GM + q] = p;
            }
        }
        for (size_t q = 0; q < in.size(); ++q) gint[gi + 4 + GM + q] = in[q];
        gint[gi + 2] = (int32_t)(uint32_t)(mask & 0xffffffffu);
        gint[gi + 3] = (int32_t)(uint32_t)(mask >> 32);
        ++ngroups;
    }
    if (C.gint.upload(gint.data(), gint.size() * 4) != QOB_OK || upload_vals(C.gcoef, gcoef, dt) != QOB_OK)
        return false;
    C.gsrc = std::move(gsrc);
    C.dev.gint = (const int32_t *)C.gint.p;
    C.dev.gcoef = C.gcoef.p;
    C.dev.ngroups = ngroups;
    C.dev.gmax = GM;
    return true;
}

static qob_status build_csc(const RowLists &Xcols, int64_t xrows, qob_dtype dt, CscTab &C) {
    // Xcols: columns of X as lists (i.e. rows of X^T)
    QOB_TRY(C.ptr.upload(Xcols.ptr.data(), Xcols.ptr.size() * 4));
    QOB_TRY(C.row.upload(Xcols.idx.data(), Xcols.idx.size() * 4));
    QOB_TRY(upload_vals(C.val, Xcols.val, dt));
    C.dev.m = xrows;
    C.dev.n = Xcols.nr;
    C.dev.colptr = (const int32_t *)C.ptr.p;
    C.dev.row = (const int32_t *)C.row.p;
    C.dev.val = C.val.p;
    int64_t bw = -1;
    const std::vector<int32_t> ord = csc_column_order(Xcols, xrows, &bw);
    C.dev.sweep_bw = (int32_t)std::min<int64_t>(bw, 1 << 30);
    if (!ord.empty()) {
        QOB_TRY(C.ord.upload(ord.data(), ord.size() * 4));
        C.dev.corder = (const int32_t *)C.ord.p;
    }
    if (g_csc_groups) build_csc_groups(Xcols, xrows, dt, C);
    return QOB_OK;
}

static qob_status refresh_sell(SparseData &S, int v, SellTab &t, qob_dtype dt, hipStream_t s);
static qob_status refresh_csc(SparseData &S, int v, CscTab &t, qob_dtype dt, hipStream_t s);

// A table of a materialised sum whose coefficients are held on the device
// takes its values from them as soon as it is built (on stream s).
static bool mat_device_coefs(const SparseData &S) { return S.mc && S.mc->dcoef_on && *S.mc->dcoef_on; }

static qob_status sparse_sell(SparseData &S, int v, qob_dtype dt, SellDev &out, hipStream_t s = nullptr) {
    std::lock_guard<std::mutex> g(S.mu);
    const int key = v * 2 + (int)dt;
    auto it = S.sell.find(key);
    if (it == S.sell.end()) {
        auto t = std::make_shared<SellTab>();
        QOB_TRY(build_sell(S.rows(v), dt, *t));
        if (mat_device_coefs(S)) QOB_TRY(refresh_sell(S, v, *t, dt, s));
        it = S.sell.emplace(key, t).first;
    }
    out = it->second->dev;
    return QOB_OK;
}

static qob_status sparse_csc(SparseData &S, int v, qob_dtype dt, CscDev &out, hipStream_t s = nullptr) {
    std::lock_guard<std::mutex> g(S.mu);
    const int key = v * 2 + (int)dt;
    auto it = S.cols.find(key);
    if (it == S.cols.end()) {
        auto t = std::make_shared<CscTab>();
        // columns of op_v(M) = rows of op_v(M)^T = rows of op_{v^T}(M)
        const RowLists cl = S.rows(v ^ V_T);
        const int64_t xr = (v & 1) ? S.n : S.m;
        QOB_TRY(build_csc(cl, xr, dt, *t));
        if (mat_device_coefs(S)) QOB_TRY(refresh_csc(S, v, *t, dt, s));
        it = S.cols.emplace(key, t).first;
    }
    out = it->second->dev;
    return QOB_OK;
}

// ---------------------------------------------------------------- dense op
struct DenseData {
    int64_t m = 0, n = 0;
    std::vector<cd> a;  // column-major m x n
    std::mutex mu;
    std::map<int, std::shared_ptr<DevBuf>> dev;  // key var*2 + dt
    // op_v(M) as a column-major host matrix (rows x cols)
    std::vector<cd> mat(int v, int64_t &rows, int64_t &cols) const {
        const bool t = v & 1, c = (v >> 1) & 1;
        rows = t ? n : m;
        cols = t ? m : n;
        std::vector<cd> r((size_t)(m * n));
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) {
                cd z = a[i + j * m];
                if (c) z = std::conj(z);
                if (t) r[j + i * n] = z;
                else r[i + j * m] = z;
            }
        return r;
    }
    RowLists rows_of(int v) const {
        int64_t r, c;
        std::vector<cd> M = mat(v, r, c);
        RowLists R;
        R.nr = r;
        R.nc = c;
        R.ptr.assign(r + 1, 0);
        for (int64_t i = 0; i < r; ++i) {
            for (int64_t j = 0; j < c; ++j) {
                R.idx.push_back((int32_t)j);
                R.val.push_back(M[i + j * r]);
            }
            R.ptr[i + 1] = (int32_t)R.idx.size();
        }
        return R;
    }
};

static qob_status dense_dev(DenseData &D, int v, qob_dtype dt, const void *&ptr) {
    std::lock_guard<std::mutex> g(D.mu);
    const int key = v * 2 + (int)dt;
    auto it = D.dev.find(key);
    if (it == D.dev.end()) {
        int64_t r, c;
        auto b = std::make_shared<DevBuf>();
        QOB_TRY(upload_vals(*b, D.mat(v, r, c), dt));
        it = D.dev.emplace(key, b).first;
    }
    ptr = it->second->p;
    return QOB_OK;
}

// ---------------------------------------------------------------- op nodes
enum Kind { K_SPARSE, K_DENSE, K_EYE, K_LT, K_LS, K_LP };

// Linear map from a LazySum's coefficients to a plan's values: slot s (a
// value of the plan's table when s < nvals, else a complex double inline in
// its program at word offset inline_off[s - nvals]) equals
// sum_p bval[p] * coef[term[p]] over p in [ptr[s], ptr[s+1]).
struct RefreshMap {
    int nslots = 0, nvals = 0;
    DevBuf ptr, term, bval, dst;
};

// Build the refresh map from the slot values obtained with each unit
// coefficient vector: cols[t][s] is slot s with coefficient e_t.
static qob_status build_refresh(const std::vector<std::vector<cd>> &cols, int nvals,
                                const std::vector<int32_t> &inline_off, RefreshMap &R) {
    const size_t nslots = (size_t)nvals + inline_off.size();
    std::vector<int32_t> ptr(1, 0), term;
    std::vector<double2> bval;
    for (size_t sl = 0; sl < nslots; ++sl) {
        for (size_t t = 0; t < cols.size(); ++t) {
            const cd b = cols[t][sl];
            if (b != cd(0, 0)) {
                term.push_back((int32_t)t);
                bval.push_back(d2(b));
            }
        }
        ptr.push_back((int32_t)term.size());
    }
    if (term.empty()) {
        term.push_back(0);
        bval.push_back(make_double2(0, 0));
    }
    R.nslots = (int)nslots;
    R.nvals = nvals;
    QOB_TRY(R.ptr.upload(ptr.data(), ptr.size() * 4));
    QOB_TRY(R.term.upload(term.data(), term.size() * 4));
    QOB_TRY(R.bval.upload(bval.data(), bval.size() * sizeof(double2)));
    std::vector<int32_t> dst = inline_off;
    if (dst.empty()) dst.push_back(0);
    QOB_TRY(R.dst.upload(dst.data(), dst.size() * 4));
    return QOB_OK;
}

static qob_status run_refresh(const RefreshMap &R, qob_dtype dt, const void *coef, void *vals, void *prog,
                              hipStream_t s) {
    hipError_t e = launch_refresh(dt, R.nslots, R.nvals, (const int32_t *)R.ptr.p, (const int32_t *)R.term.p,
                                  (const double2 *)R.bval.p, (const int32_t *)R.dst.p, (const double2 *)coef, vals,
                                  (int32_t *)prog, s);
    if (e != hipSuccess) return fail(QOB_EHIP, "coefficient refresh: %s", hipGetErrorString(e));
    return QOB_OK;
}

// Refresh map of a table of a materialised sum: slot sl holds entry src[sl]
// of the CSR (csc_space false) or CSC lists of the sum, i.e. the entry's
// coefficient-weighted contributions.  Tables of conjugated variants are not
// refreshable (the coefficients would need conjugating).
static qob_status mat_refresh_map(const MatContrib &mc, bool csc_space, const std::vector<int64_t> &src,
                                  RefreshMap &R) {
    if (src.size() >= (size_t)INT32_MAX) return fail(QOB_EUNSUPPORTED, "table too large to refresh");
    std::vector<int32_t> ptr(1, 0), term;
    std::vector<double2> bval;
    for (int64_t q : src) {
        if (q >= 0) {
            const int64_t e = csc_space ? mc.csc2csr[(size_t)q] : q;
            for (int64_t c = mc.cptr[(size_t)e]; c < mc.cptr[(size_t)e + 1]; ++c) {
                term.push_back(mc.cterm[(size_t)c]);
                bval.push_back(d2(mc.cbase[(size_t)c]));
            }
        }
        if (term.size() >= (size_t)INT32_MAX) return fail(QOB_EUNSUPPORTED, "table too large to refresh");
        ptr.push_back((int32_t)term.size());
    }
    if (term.empty()) {
        term.push_back(0);
        bval.push_back(make_double2(0, 0));
    }
    R.nslots = (int)src.size();
    R.nvals = (int)src.size();
    QOB_TRY(R.ptr.upload(ptr.data(), ptr.size() * 4));
    QOB_TRY(R.term.upload(term.data(), term.size() * 4));
    QOB_TRY(R.bval.upload(bval.data(), bval.size() * sizeof(double2)));
    const int32_t zero = 0;
    QOB_TRY(R.dst.upload(&zero, 4));
    return QOB_OK;
}

// Values of a materialised sum's table from the owner's device coefficients
// (the table was built from rows(v) / rows(v ^ V_T) of the sum).
static qob_status refresh_sell(SparseData &S, int v, SellTab &t, qob_dtype dt, hipStream_t s) {
    if (v & V_CJ) return fail(QOB_EUNSUPPORTED, "conjugated table of a sum with device-held coefficients");
    if (!t.rmap) {
        auto R = std::make_shared<RefreshMap>();
        QOB_TRY(mat_refresh_map(*S.mc, (v & V_T) != 0, t.src, *R));
        t.rmap = R;
    }
    return t.src.empty() ? QOB_OK : run_refresh(*t.rmap, dt, *S.mc->dcoef, t.val.p, nullptr, s);
}

static qob_status refresh_csc(SparseData &S, int v, CscTab &t, qob_dtype dt, hipStream_t s) {
    if (v & V_CJ) return fail(QOB_EUNSUPPORTED, "conjugated table of a sum with device-held coefficients");
    const bool csc_space = ((v ^ V_T) & V_T) != 0;
    if (!t.rmap) {
        // val holds the row lists' values in order
        std::vector<int64_t> id(S.csr.idx.size());
        for (size_t q = 0; q < id.size(); ++q) id[q] = (int64_t)q;
        auto R = std::make_shared<RefreshMap>();
        QOB_TRY(mat_refresh_map(*S.mc, csc_space, id, *R));
        t.rmap = R;
    }
    if (t.dev.colptr && !S.csr.idx.empty()) QOB_TRY(run_refresh(*t.rmap, dt, *S.mc->dcoef, t.val.p, nullptr, s));
    if (t.dev.gcoef && !t.gsrc.empty()) {
        if (!t.grmap) {
            auto R = std::make_shared<RefreshMap>();
            QOB_TRY(mat_refresh_map(*S.mc, csc_space, t.gsrc, *R));
            t.grmap = R;
        }
        QOB_TRY(run_refresh(*t.grmap, dt, *S.mc->dcoef, t.gcoef.p, nullptr, s));
    }
    return QOB_OK;
}

// every table of a materialised sum from the device coefficients (after
// they change)
static qob_status refresh_mat(SparseData &S, hipStream_t s) {
    std::lock_guard<std::mutex> g(S.mu);
    for (auto &kv : S.sell) QOB_TRY(refresh_sell(S, kv.first / 2, *kv.second, (qob_dtype)(kv.first % 2), s));
    for (auto &kv : S.cols) QOB_TRY(refresh_csc(S, kv.first / 2, *kv.second, (qob_dtype)(kv.first % 2), s));
    return QOB_OK;
}

struct GatherPlan {
    DevBuf prog, rowptr, koff, vals, site_info;
    GatherPlanDev dev;
    qob_dtype dt = QOB_C64;
    std::shared_ptr<RefreshMap> rmap;
};

struct OpNode {
    Kind kind = K_EYE;
    int64_t m = 0, n = 0;
    int var = V_N;
    std::shared_ptr<SparseData> sp;
    std::shared_ptr<DenseData> de;
    // LazyTensor
    std::vector<int64_t> shl, shr;
    cd factor = cd(1, 0);
    std::vector<int> fsite;  // 0-based, sorted
    std::vector<std::shared_ptr<OpNode>> fop;
    // LazySum
    std::vector<cd> coef;
    std::vector<std::shared_ptr<OpNode>> terms;
    // plan cache
    std::mutex mu;
    std::map<std::string, std::shared_ptr<GatherPlan>> gplans;
    std::map<std::string, std::shared_ptr<SparseData>> sums;
    std::map<std::string, std::shared_ptr<struct QubitPlan>> qplans;
    uint64_t version = 0;  // bumped when coefficients change on the host
    // LazySum coefficients held on the device (complex doubles); when
    // dcoef_on they, not coef, are current and every plan's values follow them
    DevBuf dcoef;
    bool dcoef_on = false;
};

static bool is_pure_sparse_factor(const OpNode &o) {
    return (o.kind == K_SPARSE && o.var == V_N) || o.kind == K_EYE;
}
// _is_pure_sparse (src/operators_lazytensor.jl:520) with the non-empty
// condition of :542.
static bool lt_pure_sparse(const OpNode &o) {
    if (o.kind != K_LT || o.fop.empty()) return false;
    for (auto &f : o.fop)
        if (!is_pure_sparse_factor(*f)) return false;
    return true;
}

}  // namespace qob

struct qob_op_s {
    std::shared_ptr<qob::OpNode> node;
    qob_ctx ctx = nullptr;  // the context that created the handle
};

namespace qob {

// ------------------------------------------------------------- workspace
// slots: chain temporaries, reductions, scalars, expect temporary, the
// split partial sums of a dense matrix-vector product, the widened state and
// result of a call whose state and result element types differ, op * rho of
// a variance, and two LazyProduct temporaries per nesting depth (up to 4)
enum {
    WS_TMP1 = 0, WS_TMP2 = 1, WS_RED = 2, WS_SCAL = 3, WS_EXP = 4, WS_GEMV = 5, WS_MIXB = 6, WS_MIXR = 7,
    WS_VAR = 8, WS_LP0 = 9, WS_NSLOT = 17
};
constexpr int LP_MAX_DEPTH = 4;

struct Workspace {
    DevBuf slot[WS_NSLOT];
};

// Device memory that queued work, or a live captured graph, may still read
// when the library replaces it: kept until it is safe to free.
struct Retired {
    std::vector<std::shared_ptr<void>> keep;
    std::vector<hipEvent_t> evs;  // recorded on streams that may read it
    bool need_sync = false;       // wait for a device synchronisation as well
};

}  // namespace qob

struct qob_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    qob_mode mode = QOB_MODE_FUSED;
    std::mutex mu;
    std::unordered_map<hipStream_t, std::unique_ptr<qob::Workspace>> ws;
    // captured graphs alive and memory retired while they (or queued work) may use it
    std::mutex rmu;
    int live_graphs = 0;
    std::deque<qob::Retired> retired;
    // streams created by qob_stream_create (alive until qob_stream_destroy),
    // and whether work was ever issued on a stream the library did not create
    std::unordered_set<hipStream_t> lib_streams;
    bool foreign_streams = false;
};

namespace qob {


// Free retired memory whose events have completed, in retirement order, as
// long as no captured graph of the context is alive (a graph may replay any
// of it).  synced: the device was just synchronised.
static void reap_locked(qob_ctx ctx, bool synced) {
    if (ctx->live_graphs > 0) return;
    while (!ctx->retired.empty()) {
        Retired &R = ctx->retired.front();
        bool done = synced || !R.need_sync;
        for (size_t i = 0; done && i < R.evs.size(); ++i)
            if (!synced && hipEventQuery(R.evs[i]) == hipErrorNotReady) done = false;
        if (!done) break;
        for (hipEvent_t e : R.evs) (void)hipEventDestroy(e);
        ctx->retired.pop_front();
    }
}

static bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Keep `keep` until the work queued so far that may read it has finished and
// no captured graph is alive, instead of freeing it now.  only != null: only
// work on that stream may read it (a workspace of that stream).  Otherwise
// events go on the context's stream and on the streams the library created;
// if work was ever issued on other streams (whose lifetime the library does
// not know), the memory waits for a device synchronisation instead, which
// runs at the latest every 8th such retirement.
static void retire(qob_ctx ctx, std::vector<std::shared_ptr<void>> keep, hipStream_t only = nullptr) {
    if (keep.empty()) return;
    std::vector<hipStream_t> streams;
    bool foreign;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        foreign = ctx->foreign_streams && !only;
        if (only) {
            streams.push_back(only);
        } else {
            streams.push_back(ctx->stream);
            for (hipStream_t x : ctx->lib_streams) streams.push_back(x);
        }
    }
    Retired R;
    R.keep = std::move(keep);
    R.need_sync = foreign;
    for (hipStream_t x : streams) {
        hipEvent_t e = nullptr;
        if (capturing(x) || hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            R.need_sync = true;
            continue;
        }
        if (hipEventRecord(e, x) != hipSuccess) {
            (void)hipEventDestroy(e);
            R.need_sync = true;
            continue;
        }
        R.evs.push_back(e);
    }
    std::lock_guard<std::mutex> g(ctx->rmu);
    ctx->retired.push_back(std::move(R));
    size_t waiting = 0;
    for (const Retired &x : ctx->retired) waiting += x.need_sync;
    bool synced = false;
    if (waiting >= 8 && ctx->live_graphs == 0 && !capturing(ctx->stream)) synced = hipDeviceSynchronize() == hipSuccess;
    reap_locked(ctx, synced);
}

static qob_status ws_get(qob_ctx ctx, hipStream_t s, int slot, size_t bytes, void *&p) {
    Workspace *w;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto &u = ctx->ws[s];
        if (!u) u.reset(new Workspace());
        w = u.get();
    }
    if (w->slot[slot].bytes < bytes) {
        if (capturing(s))
            return fail(QOB_EUNSUPPORTED,
                        "a workspace would grow (to %zu bytes) during stream capture: run the call once before "
                        "capturing it",
                        bytes);
        // queued work on s, or a captured graph, may still read the old
        // buffer: it is retired, not freed
        retire(ctx, {w->slot[slot].release()}, s);
        QOB_TRY(w->slot[slot].alloc(bytes));
    }
    p = w->slot[slot].p;
    return QOB_OK;
}

static hipStream_t strm(qob_ctx ctx, qob_stream s) {
    if (!s) return ctx->stream;
    const hipStream_t h = (hipStream_t)s;
    if (h != ctx->stream && !ctx->foreign_streams) {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (!ctx->lib_streams.count(h)) ctx->foreign_streams = true;
    }
    return h;
}
hipStream_t ctx_stream(qob_ctx ctx, qob_stream s) { return strm(ctx, s); }
int ctx_device(qob_ctx ctx) { return ctx->device; }

// ------------------------------------------------------------ plan building
struct TermSpec {
    cd coef;
    std::vector<int> sites;                // sorted
    std::vector<const RowLists *> X;       // X[i]: out_d x in_d matrix at sites[i]
};

struct HostPlan {
    std::vector<int32_t> prog, rowptr, koff, site_info;
    std::vector<cd> vals;
    int nterms = 0, nsites = 0;
    bool square = true;
};

static int64_t prod(const std::vector<int64_t> &v, size_t a, size_t b) {
    int64_t p = 1;
    for (size_t i = a; i < b; ++i) p *= v[i];
    return p;
}

// Compile a sum of tensor-product terms into gather tables.
//   out_shape / in_shape : per-site dims of the output and input spaces
//   A term with factor X_i (out_d x in_d) at site s_i acts as the identity
//   (truncated to min(out_d, in_d)) on every other site.
static qob_status compile_plan(const std::vector<int64_t> &out_shape, const std::vector<int64_t> &in_shape,
                               const std::vector<TermSpec> &terms, HostPlan &H) {
    const int N = (int)out_shape.size();
    H.square = (out_shape == in_shape);
    H.nsites = N;
    const int64_t Dout = prod(out_shape, 0, N), Din = prod(in_shape, 0, N);
    if (Dout >= (1ll << 31) || Din >= (1ll << 31))
        return fail(QOB_EUNSUPPORTED, "lazy gather plan needs dimensions below 2^31");
    if (!H.square && N > 32) return fail(QOB_EUNSUPPORTED, "non-square lazy operators with more than 32 sites");
    std::vector<int64_t> so(N), si(N);
    for (int s = 0; s < N; ++s) {
        so[s] = prod(out_shape, 0, s);
        si[s] = prod(in_shape, 0, s);
    }
    auto site_rec = [&](int s, int32_t lstride, int32_t tbl) {
        FastDiv a = make_fastdiv((uint32_t)so[s]), b = make_fastdiv((uint32_t)out_shape[s]);
        return std::vector<int32_t>{(int32_t)so[s], (int32_t)a.m, (int32_t)a.l, (int32_t)out_shape[s],
                                    (int32_t)b.m, (int32_t)b.l, lstride, tbl, (int32_t)si[s]};
    };
    if (!H.square) {
        for (int s = 0; s < N; ++s) {
            FastDiv a = make_fastdiv((uint32_t)so[s]), b = make_fastdiv((uint32_t)out_shape[s]);
            int32_t rec[8] = {(int32_t)so[s], (int32_t)a.m, (int32_t)a.l, (int32_t)out_shape[s],
                              (int32_t)b.m, (int32_t)b.l, (int32_t)si[s], (int32_t)in_shape[s]};
            H.site_info.insert(H.site_info.end(), rec, rec + 8);
        }
    }

    // Group terms by site set (expanded form); keep product-form terms apart.
    struct Group {
        std::vector<int> sites;
        std::vector<std::map<int64_t, cd>> rows;  // per local row: koff -> value
    };
    std::vector<Group> groups;
    std::map<std::vector<int>, size_t> gidx;
    struct Prod {
        const TermSpec *t;
    };
    std::vector<Prod> prods;

    for (const TermSpec &t : terms) {
        const int q = (int)t.sites.size();
        int64_t nloc = 1;
        bool mono = true;
        for (int i = 0; i < q; ++i) {
            const int s = t.sites[i];
            if (t.X[i]->nr != out_shape[s] || t.X[i]->nc != in_shape[s])
                return fail(QOB_EDIM, "factor at site %d is %lldx%lld, site dims are %lldx%lld", s + 1,
                            (long long)t.X[i]->nr, (long long)t.X[i]->nc, (long long)out_shape[s],
                            (long long)in_shape[s]);
            nloc *= out_shape[s];
            mono = mono && t.X[i]->monomial();
        }
        if (mono && nloc > 4096) {
            prods.push_back({&t});
            continue;
        }
        if (nloc > (1ll << 22)) return fail(QOB_EUNSUPPORTED, "LazyTensor term with local dimension %lld", (long long)nloc);
        auto it = gidx.find(t.sites);
        if (it == gidx.end()) {
            Group g;
            g.sites = t.sites;
            g.rows.resize((size_t)nloc);
            it = gidx.emplace(t.sites, groups.size()).first;
            groups.push_back(std::move(g));
        }
        Group &g = groups[it->second];
        // odometer over the local output digits (first site fastest)
        std::vector<int64_t> jd(q, 0);
        for (int64_t jloc = 0; jloc < nloc; ++jloc) {
            // Cartesian product over the factors' row entries
            std::vector<int32_t> pp(q), pe(q);
            bool empty = false;
            for (int i = 0; i < q; ++i) {
                pp[i] = t.X[i]->ptr[jd[i]];
                pe[i] = t.X[i]->ptr[jd[i] + 1];
                if (pp[i] == pe[i]) empty = true;
            }
            if (!empty) {
                std::vector<int32_t> cur = pp;
                while (true) {
                    cd v = t.coef;
                    int64_t koff = 0;
                    for (int i = 0; i < q; ++i) {
                        const int s = t.sites[i];
                        const int64_t k = t.X[i]->idx[cur[i]];
                        v *= t.X[i]->val[cur[i]];
                        koff += H.square ? (k - jd[i]) * so[s] : k * si[s];
                    }
                    g.rows[jloc][koff] += v;
                    int i = 0;
                    while (i < q) {
                        if (++cur[i] < pe[i]) break;
                        cur[i] = pp[i];
                        ++i;
                    }
                    if (i == q) break;
                }
            }
            for (int i = 0; i < q; ++i) {
                if (++jd[i] < out_shape[t.sites[i]]) break;
                jd[i] = 0;
            }
        }
    }

    // Emit expanded groups (diagonal ones first so their loads merge).
    std::stable_sort(groups.begin(), groups.end(), [&](const Group &a, const Group &b) {
        auto diag = [&](const Group &g) {
            if (!H.square) return false;
            for (auto &r : g.rows)
                for (auto &kv : r)
                    if (kv.first != 0) return false;
            return true;
        };
        return diag(a) && !diag(b);
    });
    for (const Group &g : groups) {
        bool diag = H.square;
        for (auto &r : g.rows)
            for (auto &kv : r)
                if (kv.first != 0) diag = false;
        const int q = (int)g.sites.size();
        uint32_t fmask = 0;
        for (int s : g.sites) fmask |= (s < 32) ? (1u << s) : 0u;
        const int32_t base = (int32_t)H.rowptr.size();
        H.prog.insert(H.prog.end(), {0, diag ? 1 : 0, q, base, (int32_t)fmask});
        int32_t ls = 1;
        for (int i = 0; i < q; ++i) {
            auto rec = site_rec(g.sites[i], ls, 0);
            H.prog.insert(H.prog.end(), rec.begin(), rec.end());
            ls *= (int32_t)out_shape[g.sites[i]];
        }
        for (auto &r : g.rows) {
            H.rowptr.push_back((int32_t)H.koff.size());
            for (auto &kv : r) {
                H.koff.push_back((int32_t)kv.first);
                H.vals.push_back(kv.second);
            }
        }
        H.rowptr.push_back((int32_t)H.koff.size());
        H.nterms++;
    }
    // Emit product-form terms: per site a table over the output digit.
    for (const Prod &pr : prods) {
        const TermSpec &t = *pr.t;
        const int q = (int)t.sites.size();
        uint32_t fmask = 0;
        for (int s : t.sites) fmask |= (s < 32) ? (1u << s) : 0u;
        bool diag = H.square;
        for (int i = 0; i < q; ++i)
            for (int64_t j = 0; j < t.X[i]->nr; ++j)
                for (int32_t p = t.X[i]->ptr[j]; p < t.X[i]->ptr[j + 1]; ++p)
                    if (t.X[i]->idx[p] != j) diag = false;
        H.prog.insert(H.prog.end(), {1, diag ? 1 : 0, q, 0, (int32_t)fmask});
        for (int i = 0; i < q; ++i) {
            const int s = t.sites[i];
            const int32_t tbl = (int32_t)H.koff.size();
            auto rec = site_rec(s, 0, tbl);
            H.prog.insert(H.prog.end(), rec.begin(), rec.end());
            for (int64_t j = 0; j < out_shape[s]; ++j) {
                const RowLists &X = *t.X[i];
                if (X.ptr[j + 1] > X.ptr[j]) {
                    const int64_t k = X.idx[X.ptr[j]];
                    H.koff.push_back((int32_t)(H.square ? (k - j) * so[s] : k * si[s]));
                    cd v = X.val[X.ptr[j]];
                    if (i == 0) v *= t.coef;
                    H.vals.push_back(v);
                } else {
                    H.koff.push_back(0);
                    H.vals.push_back(cd(0, 0));
                }
            }
        }
        H.nterms++;
    }
    if (H.rowptr.empty()) H.rowptr.push_back(0);
    if (H.koff.empty()) {
        H.koff.push_back(0);
        H.vals.push_back(cd(0, 0));
    }
    return QOB_OK;
}

static qob_status upload_plan(const HostPlan &H, qob_dtype dt, GatherPlan &P) {
    QOB_TRY(P.prog.upload(H.prog.data(), std::max<size_t>(H.prog.size(), 1) * 4));
    QOB_TRY(P.rowptr.upload(H.rowptr.data(), H.rowptr.size() * 4));
    QOB_TRY(P.koff.upload(H.koff.data(), H.koff.size() * 4));
    QOB_TRY(upload_vals(P.vals, H.vals, dt));
    if (!H.site_info.empty()) QOB_TRY(P.site_info.upload(H.site_info.data(), H.site_info.size() * 4));
    P.dev.prog = (const int32_t *)P.prog.p;
    P.dev.nterms = H.nterms;
    P.dev.rowptr = (const int32_t *)P.rowptr.p;
    P.dev.koff = (const int32_t *)P.koff.p;
    P.dev.vals = P.vals.p;
    P.dev.nsites = H.nsites;
    P.dev.site_info = (const int32_t *)P.site_info.p;
    P.dev.square = H.square;
    return QOB_OK;
}

// Rows of a pure-sparse factor matrix for the gather.  trans: use the
// transpose (right application / Bra).
static RowLists factor_rows(const OpNode &f, bool trans) {
    if (f.kind == K_SPARSE) return f.sp->rows(f.var ^ (trans ? V_T : V_N));
    if (f.kind == K_DENSE) return f.de->rows_of(f.var ^ (trans ? V_T : V_N));
    return trans ? identity_rows(f.n, f.m) : identity_rows(f.m, f.n);
}

// Gather plan for a set of pure-sparse LazyTensors (one term per tensor).
// trans = false: acts on a Ket / rows of rho (out space = basis_l);
// trans = true : transposed, acts on a Bra / columns of rho (out = basis_r).
static qob_status lazy_plan(OpNode &owner, const std::string &key, const std::vector<OpNode *> &lts,
                            const std::vector<cd> &coefs, bool trans, qob_dtype dt, GatherPlanDev &out,
                            bool refreshable = false, hipStream_t s = nullptr) {
    std::lock_guard<std::mutex> g(owner.mu);
    const std::string k = key + (trans ? ":T:" : ":N:") + std::to_string((int)dt);
    auto it = owner.gplans.find(k);
    if (it == owner.gplans.end()) {
        const OpNode &first = *lts[0];
        const std::vector<int64_t> &out_shape = trans ? first.shr : first.shl;
        const std::vector<int64_t> &in_shape = trans ? first.shl : first.shr;
        std::vector<std::vector<RowLists>> store(lts.size());
        std::vector<TermSpec> specs;
        for (size_t t = 0; t < lts.size(); ++t) {
            const OpNode &lt = *lts[t];
            TermSpec ts;
            ts.coef = coefs[t] * lt.factor;
            store[t].reserve(lt.fop.size());
            for (size_t i = 0; i < lt.fop.size(); ++i) {
                if (lt.fop[i]->kind == K_EYE) continue;  // identity site (truncating)
                store[t].push_back(factor_rows(*lt.fop[i], trans));
            }
            size_t k2 = 0;
            for (size_t i = 0; i < lt.fop.size(); ++i) {
                if (lt.fop[i]->kind == K_EYE) continue;
                ts.sites.push_back(lt.fsite[i]);
                ts.X.push_back(&store[t][k2++]);
            }
            specs.push_back(std::move(ts));
        }
        HostPlan H;
        QOB_TRY(compile_plan(out_shape, in_shape, specs, H));
        auto P = std::make_shared<GatherPlan>();
        QOB_TRY(upload_plan(H, dt, *P));
        P->dt = dt;
        if (refreshable) {
            // values as linear functions of the coefficients (the table
            // layout does not depend on them)
            std::vector<std::vector<cd>> cols;
            for (size_t t = 0; t < specs.size(); ++t) {
                std::vector<TermSpec> unit = specs;
                for (size_t q = 0; q < unit.size(); ++q) unit[q].coef = (q == t) ? lts[q]->factor : cd(0, 0);
                HostPlan U;
                QOB_TRY(compile_plan(out_shape, in_shape, unit, U));
                if (U.vals.size() != H.vals.size() || U.koff != H.koff || U.rowptr != H.rowptr)
                    return fail(QOB_EUNSUPPORTED, "gather plan layout depends on the coefficients");
                cols.push_back(std::move(U.vals));
            }
            P->rmap = std::make_shared<RefreshMap>();
            QOB_TRY(build_refresh(cols, (int)H.vals.size(), {}, *P->rmap));
            if (owner.dcoef_on) QOB_TRY(run_refresh(*P->rmap, dt, owner.dcoef.p, P->vals.p, nullptr, s));
        }
        it = owner.gplans.emplace(k, P).first;
    }
    out = it->second->dev;
    return QOB_OK;
}

// ------------------------------------------------------- qubit programs
// A sum of tensor-product terms on qubits whose factors are monomial 2x2
// matrices maps basis state J to J ^ F_t.  Terms are grouped by F and each
// group's weight w_F(J) = sum_t c_t prod_s v_{t,s}[bit_s(J)] is stored as a
// constant or as tables over windows of <= 8 bits (program layout in
// k_qubit.hip).  Returns false when a term does not fit this form.
int g_qubit_split = 0;  // 1: high-bit flips in a second, strided pass; 0: partner tiles read in one pass

struct QubitProgHost {
    std::vector<int32_t> prog;     // flips inside a tile (all flips when prog_hi is empty)
    std::vector<int32_t> prog_hi;  // flips of bits >= tb, for the strided second pass
    std::vector<cd> vals;
};

// tb: tile bits of the launch; states longer than a tile (and at most 2 tb
// bits) get their high-bit flips in a separate program.
static bool compile_qubit(int nbits, const std::vector<TermSpec> &terms, QubitProgHost &Q, int tb,
                          int split_flag = -1) {
    struct QT {
        cd coef;
        uint32_t F;
        std::vector<int> sites;
        std::vector<std::array<cd, 2>> v;
    };
    std::vector<QT> qts;
    for (const TermSpec &t : terms) {
        QT q;
        q.coef = t.coef;
        q.F = 0;
        bool zero = false;
        for (size_t i = 0; i < t.sites.size(); ++i) {
            const RowLists &X = *t.X[i];
            if (X.nr != 2 || X.nc != 2) return false;
            int flip = -1;
            std::array<cd, 2> v{cd(0, 0), cd(0, 0)};
            for (int j = 0; j < 2; ++j) {
                const int n = X.ptr[j + 1] - X.ptr[j];
                if (n > 1) return false;
                if (n == 1) {
                    const int k = X.idx[X.ptr[j]];
                    const int f = (k != j) ? 1 : 0;
                    if (flip >= 0 && flip != f) return false;
                    flip = f;
                    v[j] = X.val[X.ptr[j]];
                }
            }
            if (flip < 0) {
                zero = true;  // factor with no entries: the term vanishes
                break;
            }
            q.F |= (uint32_t)flip << t.sites[i];
            q.sites.push_back(t.sites[i]);
            q.v.push_back(v);
        }
        if (!zero) qts.push_back(std::move(q));
    }
    std::map<uint32_t, std::vector<const QT *>> groups;
    for (const QT &q : qts) groups[q.F].push_back(&q);
    struct GroupOut {
        uint32_t F;
        cd c;
        std::map<int, std::vector<cd>> win;
    };
    std::vector<GroupOut> outs[2];  // [0]: first pass, [1]: high-bit pass
    const bool split = (split_flag < 0 ? g_qubit_split : split_flag) && nbits > tb && nbits <= 2 * tb;
    for (auto &gq : groups) {
        GroupOut go;
        go.F = gq.first;
        go.c = cd(0, 0);
        // windows of width <= 8 starting at multiples of 6 where possible.
        // The layout depends only on the factors' entry patterns, never on
        // the coefficients, so that coefficient changes only rewrite values
        // (a term whose weight pattern is constant goes to the constant part;
        // groups are kept even when their weights cancel).
        for (const QT *q : gq.second) {
            bool flat = true;
            for (size_t i = 0; i < q->sites.size(); ++i) flat = flat && (q->v[i][0] == q->v[i][1]);
            if (q->sites.empty() || flat) {
                cd w = q->coef;
                for (size_t i = 0; i < q->sites.size(); ++i) w *= q->v[i][0];
                go.c += w;
                continue;
            }
            const int m = *std::min_element(q->sites.begin(), q->sites.end());
            const int M = *std::max_element(q->sites.begin(), q->sites.end());
            if (M - m >= 8) return false;
            int ws = (m / 6) * 6;
            if (M >= ws + 8) ws = m;
            const int nb = std::min(8, nbits - ws);
            auto &tab = go.win[ws];
            if (tab.empty()) tab.assign((size_t)1 << nb, cd(0, 0));
            for (int idx = 0; idx < (1 << nb); ++idx) {
                cd w = q->coef;
                for (size_t i = 0; i < q->sites.size(); ++i) w *= q->v[i][(idx >> (q->sites[i] - ws)) & 1];
                tab[idx] += w;
            }
        }
        outs[(split && (go.F >> tb) != 0) ? 1 : 0].push_back(std::move(go));
    }
    auto put_cd = [](std::vector<int32_t> &P, cd z) {
        int32_t w[4];
        const double d[2] = {z.real(), z.imag()};
        std::memcpy(w, d, sizeof(w));
        P.insert(P.end(), w, w + 4);
    };
    Q.prog.clear();
    Q.prog_hi.clear();
    Q.vals.clear();
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<GroupOut> &gs = outs[pass];
        if (pass == 1 && gs.empty()) break;
        std::vector<int32_t> &P = pass ? Q.prog_hi : Q.prog;
        // constant-weight groups first (fixed-size records), then window groups
        std::stable_partition(gs.begin(), gs.end(), [](const GroupOut &g) { return g.win.empty(); });
        int nconst = 0;
        for (auto &g : gs) nconst += g.win.empty() ? 1 : 0;
        // header word 2: the first constant group whose flip leaves a tile of
        // 2^min(tb, nbits) states (constant groups are in ascending F, so the
        // ones before it read the tile and the rest read partner tiles)
        int first_out = 0;
        const int tbe = std::min(tb, nbits);
        while (first_out < nconst && (gs[first_out].F >> tbe) == 0) ++first_out;
        P.insert(P.end(), {nconst, (int32_t)gs.size() - nconst, pass == 0 ? first_out : 0, 0});
        for (auto &g : gs) {
            P.insert(P.end(), {(int32_t)g.F, (int32_t)g.win.size(), 0, 0});
            put_cd(P, g.c);
            for (auto &w : g.win) {
                const int nb = std::min(8, nbits - w.first);
                P.insert(P.end(), {w.first, nb, (int32_t)Q.vals.size(), 0});
                Q.vals.insert(Q.vals.end(), w.second.begin(), w.second.end());
            }
        }
    }
    if (Q.vals.empty()) Q.vals.push_back(cd(0, 0));
    return true;
}

struct QubitPlan {
    bool ok = false;
    int nbits = 0;
    int tbits = 12;
    int32_t hi_off = -1;  // offset of the high-bit program in prog, or -1
    DevBuf prog, vals;
    qob_dtype dt = QOB_C64;
    std::shared_ptr<RefreshMap> rmap;
};

// Word offsets of the complex doubles held inline in a qubit program (the
// constant weights), program starting at word base.
static void qubit_inline_offsets(const std::vector<int32_t> &p, int32_t base, std::vector<int32_t> &off) {
    if (p.empty()) return;
    const int nconst = p[0], nwg = p[1];
    size_t i = 4;
    for (int g = 0; g < nconst; ++g, i += 8) off.push_back(base + (int32_t)i + 4);
    for (int g = 0; g < nwg; ++g) {
        const int nwin = p[i + 1];
        off.push_back(base + (int32_t)i + 4);
        i += 8 + 4 * (size_t)nwin;
    }
}

static cd inline_value(const std::vector<int32_t> &p, int32_t off) {
    double d[2];
    std::memcpy(d, &p[(size_t)off], sizeof(d));
    return cd(d[0], d[1]);
}


// Shift the table offsets of a compiled qubit program by vb.
static void rebase_qubit_prog(std::vector<int32_t> &p, int32_t vb) {
    if (p.empty()) return;
    const int nconst = p[0], nwg = p[1];
    size_t i = 4 + 8 * (size_t)nconst;
    for (int g = 0; g < nwg; ++g) {
        const int nwin = p[i + 1];
        i += 8;
        for (int k = 0; k < nwin; ++k, i += 4) p[i + 2] += vb;
    }
}

// Terms of a LazySum / LazyTensor as qubit terms (trans: transposed factors).
static bool qubit_terms(OpNode &op, bool trans, std::vector<TermSpec> &specs, std::vector<std::vector<RowLists>> &store,
                        int &nbits, std::vector<cd> *base = nullptr) {
    std::vector<OpNode *> lts;
    std::vector<cd> coefs;
    if (op.kind == K_LT) {
        lts.push_back(&op);
        coefs.push_back(cd(1, 0));
    } else if (op.kind == K_LS) {
        for (size_t t = 0; t < op.terms.size(); ++t) {
            lts.push_back(op.terms[t].get());
            coefs.push_back(op.coef[t]);
        }
    } else {
        return false;
    }
    if (lts.empty()) return false;
    for (OpNode *lt : lts) {
        if (!lt_pure_sparse(*lt) || lt->shl != lt->shr) return false;
        for (int64_t d : lt->shl)
            if (d != 2) return false;
    }
    nbits = (int)lts[0]->shl.size();
    if (nbits < 1 || nbits > 30) return false;
    store.assign(lts.size(), {});
    for (size_t t = 0; t < lts.size(); ++t) {
        const OpNode &lt = *lts[t];
        TermSpec ts;
        ts.coef = coefs[t] * lt.factor;
        if (base) base->push_back(lt.factor);
        for (size_t i = 0; i < lt.fop.size(); ++i)
            if (lt.fop[i]->kind != K_EYE) store[t].push_back(factor_rows(*lt.fop[i], trans));
        size_t k2 = 0;
        for (size_t i = 0; i < lt.fop.size(); ++i) {
            if (lt.fop[i]->kind == K_EYE) continue;
            ts.sites.push_back(lt.fsite[i]);
            ts.X.push_back(&store[t][k2++]);
        }
        specs.push_back(std::move(ts));
    }
    return true;
}

int g_qubit_enable = 1;
extern int g_qubit_tbits;
int qubit_tbits(qob_dtype dt);

static qob_status qubit_plan(OpNode &op, bool trans, qob_dtype dt, hipStream_t s, std::shared_ptr<QubitPlan> &out,
                             bool nosplit = false) {
    std::lock_guard<std::mutex> g(op.mu);
    const int tbd = qubit_tbits(dt);
    const int split = nosplit ? 0 : g_qubit_split;
    const std::string k = "qb:" + std::to_string(op.version) + (trans ? ":T:" : ":N:") + std::to_string((int)dt) +
                          ":" + std::to_string(tbd) + ":" + std::to_string(split);
    auto it = op.qplans.find(k);
    if (it == op.qplans.end()) {
        auto P = std::make_shared<QubitPlan>();
        std::vector<TermSpec> specs;
        std::vector<std::vector<RowLists>> store;
        std::vector<cd> base;
        int nbits = 0;
        QubitProgHost H;
        // a state smaller than the tile gets a tile of its own size, at least
        // 2^11 (one workgroup with half or more of its lanes busy)
        const bool terms = qubit_terms(op, trans, specs, store, nbits, &base);
        const int tb = nbits < tbd ? std::max(std::min(tbd, 11), nbits) : tbd;
        if (terms && compile_qubit(nbits, specs, H, tb, split)) {
            std::vector<int32_t> all = H.prog;
            if (!H.prog_hi.empty()) {
                P->hi_off = (int32_t)all.size();
                all.insert(all.end(), H.prog_hi.begin(), H.prog_hi.end());
            }
            QOB_TRY(P->prog.upload(all.data(), all.size() * 4));
            QOB_TRY(upload_vals(P->vals, H.vals, dt));
            P->nbits = nbits;
            P->tbits = tb;
            P->dt = dt;
            P->ok = true;
            if (op.kind == K_LS) {
                // values as linear functions of the coefficients: compile
                // once per unit coefficient vector (same layout every time)
                std::vector<int32_t> off;
                qubit_inline_offsets(H.prog, 0, off);
                if (!H.prog_hi.empty()) qubit_inline_offsets(H.prog_hi, P->hi_off, off);
                std::vector<std::vector<cd>> cols;
                bool same = true;
                for (size_t t = 0; t < specs.size() && same; ++t) {
                    std::vector<TermSpec> unit = specs;
                    for (size_t q = 0; q < unit.size(); ++q) unit[q].coef = (q == t) ? base[q] : cd(0, 0);
                    QubitProgHost U;
                    same = compile_qubit(nbits, unit, U, tb, split) && U.prog.size() == H.prog.size() &&
                           U.prog_hi.size() == H.prog_hi.size() && U.vals.size() == H.vals.size();
                    if (!same) break;
                    std::vector<int32_t> ua = U.prog;
                    ua.insert(ua.end(), U.prog_hi.begin(), U.prog_hi.end());
                    std::vector<cd> col = U.vals;
                    for (int32_t o : off) col.push_back(inline_value(ua, o));
                    cols.push_back(std::move(col));
                }
                if (same) {
                    P->rmap = std::make_shared<RefreshMap>();
                    QOB_TRY(build_refresh(cols, (int)H.vals.size(), off, *P->rmap));
                    if (op.dcoef_on) QOB_TRY(run_refresh(*P->rmap, dt, op.dcoef.p, P->vals.p, P->prog.p, s));
                } else if (op.dcoef_on) {
                    return fail(QOB_EUNSUPPORTED, "qubit plan values are not linear in the coefficients");
                }
            }
        }
        it = op.qplans.emplace(k, P).first;
    }
    out = it->second;
    return QOB_OK;
}

// ------------------------------------------------------------ apply engine
// A "view" of a dense state for one application: data of size lead x D x
// trail where the operator acts on the middle index D (decomposed by the
// operator's site shape).  Ket: lead = trail = 1.  rho left: lead = 1,
// trail = columns.  rho right: lead = rows, trail = 1.  Bra: as a Ket of the
// transposed operator.
struct Apply {
    qob_ctx ctx;
    hipStream_t s;
    qob_dtype dt;
    int64_t lead, trail;  // one of them is 1
    bool trans;           // use transposed factors (right / Bra)
    int depth = 0;        // LazyProduct nesting (selects its temporaries)
};

static qob_status apply_op(const Apply &ap, OpNode &op, const void *in, void *out, cd alpha, cd beta);

static int64_t op_out_dim(const OpNode &op, bool trans) { return trans ? op.n : op.m; }
static int64_t op_in_dim(const OpNode &op, bool trans) { return trans ? op.m : op.n; }

// out = alpha * in (padded/truncated) + beta * out over the lead/trail view
static qob_status apply_embed(const Apply &ap, int64_t dout, int64_t din, const void *in, void *out, cd alpha,
                              cd beta) {
    hipError_t e;
    if (ap.lead == 1) e = launch_embed(ap.dt, in, din, ap.trail, out, dout, ap.trail, d2(alpha), d2(beta), ap.s);
    else e = launch_embed(ap.dt, in, ap.lead, din, out, ap.lead, dout, d2(alpha), d2(beta), ap.s);
    if (e != hipSuccess) return fail(QOB_EHIP, "embed launch: %s", hipGetErrorString(e));
    return QOB_OK;
}

static qob_status apply_scale(const Apply &ap, int64_t dout, void *out, cd beta) {
    hipError_t e = launch_scale(ap.dt, dout * ap.lead * ap.trail, d2(beta), out, ap.s);
    if (e != hipSuccess) return fail(QOB_EHIP, "scale launch: %s", hipGetErrorString(e));
    return QOB_OK;
}

// Run a compiled gather plan on the view.
static qob_status apply_gather(const Apply &ap, const GatherPlanDev &P, int64_t din, int64_t dout, const void *in,
                               void *out, cd alpha, cd beta) {
    hipError_t e;
    if (ap.lead == 1) e = launch_gather_rows(ap.dt, P, in, din, out, dout, ap.trail, d2(alpha), d2(beta), ap.s);
    else e = launch_gather_cols(ap.dt, P, in, din, out, dout, ap.lead, d2(alpha), d2(beta), ap.s);
    if (e != hipSuccess) return fail(QOB_EHIP, "gather launch: %s", hipGetErrorString(e));
    return QOB_OK;
}

// SparseOperator (src/operators_sparse.jl:199-202 -> src/sparsematrix.jl)
static qob_status apply_sparse(const Apply &ap, OpNode &op, const void *in, void *out, cd alpha, cd beta) {
    const int v = op.var ^ (ap.trans ? V_T : V_N);
    if (ap.lead == 1) {
        SellDev X;
        QOB_TRY(sparse_sell(*op.sp, v, ap.dt, X, ap.s));
        hipError_t e = launch_sell_left(ap.dt, X, in, ap.trail, out, d2(alpha), d2(beta), ap.s);
        if (e != hipSuccess) return fail(QOB_EHIP, "sparse left launch: %s", hipGetErrorString(e));
    } else {
        // right application B * op: columns of op_v^T ... the view applies
        // X = op_v along the column index: out[:, c] = sum_r X[c, r] in[:, r]
        // i.e. R = B * X^T  -> CSC of X^T = rows of X
        CscDev C;
        QOB_TRY(sparse_csc(*op.sp, v ^ V_T, ap.dt, C, ap.s));
        hipError_t e = launch_csc_right(ap.dt, C, in, ap.lead, out, d2(alpha), d2(beta), ap.s);
        if (e != hipSuccess) return fail(QOB_EHIP, "sparse right launch: %s", hipGetErrorString(e));
    }
    return QOB_OK;
}

// DenseOperator (src/operators_dense.jl:394-396, BLAS gemm / gemv there):
// a mode product over the whole view with the variant of the matrix the
// application needs (op_v, or its transpose for a Bra or a right
// application) as the factor.
static qob_status run_mode(const Apply &ap, const void *F, int64_t d_out, int64_t d_in, const void *in, void *out,
                           int64_t n1, int64_t n3, cd alpha, cd beta) {
    void *ws = nullptr;
    const size_t wb = mode_workspace_bytes(d_out, d_in, n1, n3);
    if (wb) QOB_TRY(ws_get(ap.ctx, ap.s, WS_GEMV, wb, ws));
    hipError_t e = launch_mode(ap.dt, F, d_out, d_in, in, out, n1, n3, d2(alpha), d2(beta), ap.s, ws);
    if (e != hipSuccess) return fail(QOB_EHIP, "dense mode-product launch: %s", hipGetErrorString(e));
    return QOB_OK;
}

static qob_status apply_dense(const Apply &ap, OpNode &op, const void *in, void *out, cd alpha, cd beta) {
    const void *F;
    QOB_TRY(dense_dev(*op.de, op.var ^ (ap.trans ? V_T : V_N), ap.dt, F));
    return run_mode(ap, F, op_out_dim(op, ap.trans), op_in_dim(op, ap.trans), in, out, ap.lead, ap.trail, alpha,
                    beta);
}

// One step of the tensor-product chain: mode product on site k of the
// current shape (between lead and trail).
struct ChainOp {
    int site;
    OpNode *f;            // factor node (dense / sparse / eye)
    int64_t d_out, d_in;  // effective dims of the (possibly transposed) factor
};

static qob_status chain_step(const Apply &ap, OpNode &owner, const std::string &key, const ChainOp &c,
                             const std::vector<int64_t> &shape_in, const void *in, void *out, cd alpha, cd beta) {
    const int N = (int)shape_in.size();
    int64_t n1 = ap.lead * prod(shape_in, 0, c.site);
    int64_t n3 = prod(shape_in, c.site + 1, N) * ap.trail;
    OpNode &f = *c.f;
    if (f.kind == K_DENSE) {
        const void *A;
        QOB_TRY(dense_dev(*f.de, f.var ^ (ap.trans ? V_T : V_N), ap.dt, A));  // d_out x d_in
        return run_mode(ap, A, c.d_out, c.d_in, in, out, n1, n3, alpha, beta);
    }
    // sparse factor or isometry: one-term gather over the full view
    std::vector<int64_t> full_in, full_out;
    full_in.push_back(n1);
    full_in.push_back(c.d_in);
    full_in.push_back(n3);
    full_out = full_in;
    full_out[1] = c.d_out;
    GatherPlanDev P;
    {
        std::lock_guard<std::mutex> g(owner.mu);
        char kb[160];
        snprintf(kb, sizeof kb, "%s:chain:%d:%lld:%lld:%lld:%lld:%d:%d", key.c_str(), c.site, (long long)n1,
                 (long long)c.d_in, (long long)c.d_out, (long long)n3, (int)ap.trans, (int)ap.dt);
        auto it = owner.gplans.find(kb);
        if (it == owner.gplans.end()) {
            RowLists X = factor_rows(f, ap.trans);
            TermSpec ts;
            ts.coef = cd(1, 0);
            ts.sites = {1};
            ts.X = {&X};
            HostPlan H;
            QOB_TRY(compile_plan(full_out, full_in, {ts}, H));
            auto GP = std::make_shared<GatherPlan>();
            QOB_TRY(upload_plan(H, ap.dt, *GP));
            it = owner.gplans.emplace(kb, GP).first;
        }
        P = it->second->dev;
    }
    hipError_t e = launch_gather_rows(ap.dt, P, in, n1 * c.d_in * n3, out, n1 * c.d_out * n3, 1, d2(alpha), d2(beta),
                                      ap.s);
    if (e != hipSuccess) return fail(QOB_EHIP, "gather launch: %s", hipGetErrorString(e));
    return QOB_OK;
}

// LazyTensor through the tensor-product chain: _tp_sum_matmul!
// (src/operators_lazytensor.jl:443-488) with the operator list of
// _tpops_tuple (:523-535, square identities dropped) followed by the
// explicit isometries of _explicit_isometries (:491-514).
static qob_status apply_lt_chain(const Apply &ap, OpNode &op, const void *in, void *out, cd alpha, cd beta) {
    const std::vector<int64_t> &sh_out = ap.trans ? op.shr : op.shl;
    const std::vector<int64_t> &sh_in = ap.trans ? op.shl : op.shr;
    const int N = (int)sh_in.size();
    std::vector<ChainOp> ops;
    std::vector<bool> used(N, false);
    for (size_t i = 0; i < op.fop.size(); ++i) {
        OpNode &f = *op.fop[i];
        used[op.fsite[i]] = true;
        if (f.kind == K_EYE && f.m == f.n) continue;  // square Eye: identity
        const int64_t dout = ap.trans ? f.n : f.m, din = ap.trans ? f.m : f.n;
        ops.push_back({op.fsite[i], &f, dout, din});
    }
    // isometries for sites whose dims differ and that carry no factor
    std::vector<std::shared_ptr<OpNode>> isos;
    for (int s = 0; s < N; ++s) {
        if (sh_out[s] != sh_in[s] && !used[s]) {
            auto e = std::make_shared<OpNode>();
            e->kind = K_EYE;
            e->m = ap.trans ? sh_in[s] : sh_out[s];
            e->n = ap.trans ? sh_out[s] : sh_in[s];
            isos.push_back(e);
            ops.push_back({s, e.get(), sh_out[s], sh_in[s]});
        }
    }
    const cd a = alpha * op.factor;
    const int64_t Din = prod(sh_in, 0, N), Dout = prod(sh_out, 0, N);
    // All factors dense and square on a column that fits in LDS: one fused
    // MFMA pass over the state instead of one pass per factor.
    if (g_fused_lt && ap.lead == 1 && !ops.empty() && sh_in == sh_out && Din < (1 << 30)) {
        FusedModes FM;
        FM.D = (int)Din;
        bool ok = (int)ops.size() <= FUSED_MAXMODES;
        for (size_t i = 0; ok && i < ops.size(); ++i) {
            const ChainOp &c = ops[i];
            if (c.f->kind != K_DENSE || c.d_in != c.d_out) {
                ok = false;
                break;
            }
            FM.n1[i] = (int)prod(sh_in, 0, c.site);
            FM.d[i] = (int)c.d_in;
            QOB_TRY(dense_dev(*c.f->de, c.f->var ^ (ap.trans ? V_T : V_N), ap.dt, FM.A[i]));
        }
        FM.nmodes = (int)ops.size();
        if (ok) {
            // factors on sites 1, 2 (and 3): modes 1 and 2 chained in registers
            hipError_t e = launch_lt_reg(ap.dt, FM, in, out, ap.trail, d2(a), d2(beta), ap.s);
            if (e == hipSuccess) return QOB_OK;
            if (e != hipErrorNotSupported) return fail(QOB_EHIP, "LazyTensor register-chain launch: %s", hipGetErrorString(e));
            (void)hipGetLastError();
            // no factor on a first site of dimension 16, factors on sites 2
            // and 3: a column is a tile of 16 "rows" with leading dimension 16
            e = launch_lt_reg_tile(ap.dt, FM, in, out, ap.trail, 1, 16, d2(a), d2(beta), ap.s);
            if (e == hipSuccess) return QOB_OK;
            if (e != hipErrorNotSupported)
                return fail(QOB_EHIP, "LazyTensor register-chain tile launch: %s", hipGetErrorString(e));
            (void)hipGetLastError();
        }
        if (ok && fused_supported(ap.dt, FM)) {
            hipError_t e = launch_lt_fused(ap.dt, FM, in, out, ap.trail, d2(a), d2(beta), ap.s);
            if (e != hipSuccess) return fail(QOB_EHIP, "fused LazyTensor launch: %s", hipGetErrorString(e));
            return QOB_OK;
        }
    }
    // Right application on a column-major matrix with enough rows: the
    // factors of the leading sites 0..k-1 (dimensions multiplying to P, with
    // 16 P elements per tile fitting the kernel) run in one fused pass over
    // tiles of 16 rows x P columns; the factors of the other sites run first
    // through the chain into a temporary (factors on different sites
    // commute).  Each pass reads and writes the matrix once.
    if (g_fused_lt && g_fused_tiles && ap.lead >= 64 && ap.lead % 16 == 0 && ap.trail == 1 && !ops.empty() &&
        sh_in == sh_out && Din < (1 << 30)) {
        bool dense = true;
        for (const ChainOp &c : ops)
            if (c.f->kind != K_DENSE || c.d_in != c.d_out) dense = false;
        const int64_t dmax = ap.dt == QOB_C64 ? 8192 : 16384;
        int k = 0;
        int64_t P = 1;
        while (k < N && 16 * P * sh_in[k] <= dmax) P *= sh_in[k++];
        std::vector<size_t> tile_i, rest_i;
        for (size_t i = 0; i < ops.size(); ++i) (ops[i].site < k ? tile_i : rest_i).push_back(i);
        if (dense && !tile_i.empty() && (int)tile_i.size() <= FUSED_MAXMODES) {
            FusedModes FM;
            FM.D = (int)(16 * P);
            FM.nmodes = (int)tile_i.size();
            for (size_t t = 0; t < tile_i.size(); ++t) {
                const ChainOp &c = ops[tile_i[t]];
                FM.n1[t] = (int)(16 * prod(sh_in, 0, c.site));
                FM.d[t] = (int)c.d_in;
                QOB_TRY(dense_dev(*c.f->de, c.f->var ^ (ap.trans ? V_T : V_N), ap.dt, FM.A[t]));
            }
            if (fused_supported(ap.dt, FM)) {
                const void *src = in;
                if (!rest_i.empty()) {
                    std::vector<int64_t> cur = sh_in;
                    const size_t el = (size_t)(Din * ap.lead);
                    void *t1, *t2 = nullptr;
                    QOB_TRY(ws_get(ap.ctx, ap.s, WS_TMP1, el * esize(ap.dt), t1));
                    if (rest_i.size() > 1) QOB_TRY(ws_get(ap.ctx, ap.s, WS_TMP2, el * esize(ap.dt), t2));
                    QOB_TRY(chain_step(ap, op, "lt", ops[rest_i[0]], cur, in, t1, cd(1, 0), cd(0, 0)));
                    for (size_t i = 1; i < rest_i.size(); ++i) {
                        QOB_TRY(chain_step(ap, op, "lt", ops[rest_i[i]], cur, t1, t2, cd(1, 0), cd(0, 0)));
                        std::swap(t1, t2);
                    }
                    src = t1;
                }
                const int64_t nrb = ap.lead / 16;
                // two factors filling the tile: chained in registers
                hipError_t e = launch_lt_reg_tile(ap.dt, FM, src, out, nrb * (Din / P), nrb, ap.lead, d2(a),
                                                  d2(beta), ap.s);
                if (e == hipSuccess) return QOB_OK;
                if (e != hipErrorNotSupported)
                    return fail(QOB_EHIP, "LazyTensor register-chain tile launch: %s", hipGetErrorString(e));
                (void)hipGetLastError();
                e = launch_lt_fused(ap.dt, FM, src, out, nrb * (Din / P), d2(a), d2(beta), ap.s, ap.lead, nrb);
                if (e != hipSuccess) return fail(QOB_EHIP, "fused LazyTensor tile launch: %s", hipGetErrorString(e));
                return QOB_OK;
            }
        }
    }
    if (ops.empty()) {
        // n_ops == 0: result = alpha * b + beta * result (:453-458)
        if (Din != Dout) return fail(QOB_EDIM, "LazyTensor without factors on unequal dimensions");
        return apply_embed(ap, Dout, Din, in, out, a, beta);
    }
    // shapes after each step
    std::vector<int64_t> cur = sh_in;
    if (ops.size() == 1) return chain_step(ap, op, "lt", ops[0], cur, in, out, a, beta);
    // temporaries sized for the largest intermediate
    size_t maxel = 0;
    {
        std::vector<int64_t> c2 = sh_in;
        for (size_t i = 0; i + 1 < ops.size(); ++i) {
            c2[ops[i].site] = ops[i].d_out;
            maxel = std::max<size_t>(maxel, (size_t)(prod(c2, 0, N) * ap.lead * ap.trail));
        }
    }
    void *t1, *t2;
    QOB_TRY(ws_get(ap.ctx, ap.s, WS_TMP1, maxel * esize(ap.dt), t1));
    if (ops.size() > 2) QOB_TRY(ws_get(ap.ctx, ap.s, WS_TMP2, maxel * esize(ap.dt), t2));
    else t2 = nullptr;
    QOB_TRY(chain_step(ap, op, "lt", ops[0], cur, in, t1, a, cd(0, 0)));
    cur[ops[0].site] = ops[0].d_out;
    for (size_t i = 1; i + 1 < ops.size(); ++i) {
        QOB_TRY(chain_step(ap, op, "lt", ops[i], cur, t1, t2, cd(1, 0), cd(0, 0)));
        cur[ops[i].site] = ops[i].d_out;
        std::swap(t1, t2);
    }
    return chain_step(ap, op, "lt", ops.back(), cur, t1, out, cd(1, 0), beta);
}

// Qubit Pauli-type sums on Kets / columns of rho (k_qubit.hip).  Sets done
// when the operator qualified and the launch was issued.
static qob_status apply_qubit(const Apply &ap, OpNode &op, const void *in, void *out, cd alpha, cd beta, bool &done) {
    done = false;
    if (!g_qubit_enable || ap.lead != 1) return QOB_OK;
    std::shared_ptr<QubitPlan> P;
    QOB_TRY(qubit_plan(op, ap.trans, ap.dt, ap.s, P));
    if (!P->ok) return QOB_OK;
    QubitStateDev S;
    S.nbits = P->nbits;
    S.prog_off = 0;
    S.prog_hi_off = P->hi_off;
    S.in = in;
    S.out = out;
    hipError_t e = hipSuccess;
    if (P->hi_off >= 0)  // high-bit flips first: out = alpha * high + beta * out
        e = launch_qubit_hi_cols(ap.dt, S, ap.trail, P->tbits, (const int32_t *)P->prog.p, P->vals.p, d2(alpha),
                                 d2(beta), ap.s);
    if (e == hipSuccess)
        e = launch_qubit_cols(ap.dt, S, ap.trail, P->tbits, (const int32_t *)P->prog.p, P->vals.p, d2(alpha),
                              d2(beta), make_double2(0, 0), ap.s);
    if (e != hipSuccess) return fail(QOB_EHIP, "qubit launch: %s", hipGetErrorString(e));
    done = true;
    return QOB_OK;
}

extern int g_ls_materialize;
static bool ls_materialize(OpNode &op, const std::vector<OpNode *> &lts, const std::vector<cd> &lcoef,
                           const std::vector<size_t> &lt_terms, const std::vector<size_t> &sp_terms,
                           std::shared_ptr<SparseData> &out);

static qob_status apply_lt(const Apply &ap, OpNode &op, const void *in, void *out, cd alpha, cd beta) {
    const int64_t dout = op_out_dim(op, ap.trans);
    if (is_zero(alpha)) return apply_scale(ap, dout, out, beta);  // :540
    bool done = false;
    QOB_TRY(apply_qubit(ap, op, in, out, alpha, beta, done));
    if (done) return QOB_OK;
    const int64_t nvec = ap.lead == 1 ? ap.trail : ap.lead;
    if (lt_pure_sparse(op) && g_ls_materialize && nvec >= 8 && op.m < (1ll << 31) && op.n < (1ll << 31)) {
        // on a density matrix: the tensor's sparse matrix (see ls_materialize)
        std::shared_ptr<SparseData> SD;
        if (ls_materialize(op, {&op}, {cd(1, 0)}, {0}, {}, SD)) {
            OpNode tmp;
            tmp.kind = K_SPARSE;
            tmp.m = op.m;
            tmp.n = op.n;
            tmp.sp = SD;
            return apply_sparse(ap, tmp, in, out, alpha, beta);
        }
    }
    if (lt_pure_sparse(op)) {
        // _mul_puresparse! -> _gemm_puresparse (:729-751)
        GatherPlanDev P;
        std::vector<OpNode *> one{&op};
        QOB_TRY(lazy_plan(op, "lt", one, {cd(1, 0)}, ap.trans, ap.dt, P));
        return apply_gather(ap, P, op_in_dim(op, ap.trans), dout, in, out, alpha, beta);
    }
    return apply_lt_chain(ap, op, in, out, alpha, beta);
}

// Materialised LazySum.  On a density matrix (eight or more columns, or rows
// for the right application) an all-sparse LazySum is cheaper as its summed
// sparse matrix: the table is read once per launch and the sparse kernels
// (closed blocks, component groups) then read rho once, where the gather
// decodes every term's multi-index per element.  The sum of the all-sparse
// LazyTensor terms (coefficient and factor folded in; sites without a factor
// are identities truncated to min(left, right) dimension, as in the gather)
// and the SparseOperator terms is expanded on the host, duplicate columns of
// a row summed, and cached per coefficient version.  Not used with
// device-held coefficients (their values change on the device) or above
// 2^22 stored entries.
int g_ls_materialize = 1;

static bool ls_materialize(OpNode &op, const std::vector<OpNode *> &lts, const std::vector<cd> &lcoef,
                           const std::vector<size_t> &lt_terms, const std::vector<size_t> &sp_terms,
                           std::shared_ptr<SparseData> &out) {
    std::lock_guard<std::mutex> g(op.mu);
    const std::string key = "mat" + std::to_string(op.version);
    auto it = op.sums.find(key);
    if (it != op.sums.end()) {
        out = it->second;
        return out != nullptr;
    }
    out = nullptr;
    const std::vector<int64_t> &shl = lts[0]->shl, &shr = lts[0]->shr;
    const int ns = (int)shl.size();
    // stored entries: per term, product over sites of the factor's nnz (or the identity's size)
    double est = 0;
    std::vector<std::vector<RowLists>> store(lts.size());
    std::vector<TermSpec> specs;
    for (size_t t = 0; t < lts.size(); ++t) {
        const OpNode &lt = *lts[t];
        if (lt.shl != shl || lt.shr != shr) {
            op.sums.emplace(key, nullptr);
            return false;
        }
        TermSpec ts;
        ts.coef = lt.factor;  // the coefficient lcoef[t] is applied when entries are summed
        for (size_t i = 0; i < lt.fop.size(); ++i)
            if (lt.fop[i]->kind != K_EYE) store[t].push_back(factor_rows(*lt.fop[i], false));
        size_t k2 = 0;
        for (size_t i = 0; i < lt.fop.size(); ++i) {
            if (lt.fop[i]->kind == K_EYE) continue;
            ts.sites.push_back(lt.fsite[i]);
            ts.X.push_back(&store[t][k2++]);
        }
        double e = 1;
        for (int site = 0; site < ns; ++site) {
            auto f = std::find(ts.sites.begin(), ts.sites.end(), site);
            e *= (f == ts.sites.end()) ? (double)std::min(shl[site], shr[site])
                                       : (double)ts.X[f - ts.sites.begin()]->idx.size();
        }
        est += e;
        specs.push_back(std::move(ts));
    }
    for (size_t t : sp_terms) est += (double)op.terms[t]->sp->csr.idx.size();
    if (est > (double)(1 << 22)) {
        op.sums.emplace(key, nullptr);
        return false;
    }
    const int64_t m = op.m, n = op.n;
    std::vector<int64_t> stl(ns, 1), str(ns, 1);
    for (int site = 1; site < ns; ++site) {
        stl[site] = stl[site - 1] * shl[site - 1];
        str[site] = str[site - 1] * shr[site - 1];
    }
    std::vector<RowLists> sprows;
    for (size_t t : sp_terms) sprows.push_back(op.terms[t]->sp->rows(op.terms[t]->var));
    auto S = std::make_shared<SparseData>();
    S->m = m;
    S->n = n;
    S->csr.nr = m;
    S->csr.nc = n;
    S->csr.ptr.assign((size_t)m + 1, 0);
    std::vector<int64_t> dig(ns);
    // (column, term, value without the term's coefficient)
    struct Contrib {
        int64_t k;
        int32_t t;
        cd b;
    };
    std::vector<Contrib> row;
    auto MC = std::make_shared<MatContrib>();
    MC->cptr.push_back(0);
    // coefficient of term t (a single LazyTensor: one term of coefficient 1)
    const std::vector<cd> cf = op.kind == K_LS ? op.coef : std::vector<cd>(1, cd(1, 0));
    for (int64_t i = 0; i < m; ++i) {
        row.clear();
        int64_t rem = i;
        for (int site = ns - 1; site >= 0; --site) {
            dig[site] = rem / stl[site];
            rem %= stl[site];
        }
        for (size_t st = 0; st < specs.size(); ++st) {
            const TermSpec &ts = specs[st];
            const int32_t term = (int32_t)lt_terms[st];
            // depth-first over sites: column offset and value so far
            std::function<void(int, int64_t, cd)> rec = [&](int site, int64_t k, cd v) {
                if (site == ns) {
                    row.push_back({k, term, v});
                    return;
                }
                auto f = std::find(ts.sites.begin(), ts.sites.end(), site);
                if (f == ts.sites.end()) {
                    if (dig[site] < shr[site]) rec(site + 1, k + dig[site] * str[site], v);
                    return;
                }
                const RowLists &X = *ts.X[f - ts.sites.begin()];
                for (int32_t p = X.ptr[dig[site]]; p < X.ptr[dig[site] + 1]; ++p)
                    rec(site + 1, k + (int64_t)X.idx[p] * str[site], v * X.val[p]);
            };
            rec(0, 0, ts.coef);
        }
        for (size_t q = 0; q < sp_terms.size(); ++q) {
            const RowLists &R = sprows[q];
            for (int32_t p = R.ptr[i]; p < R.ptr[i + 1]; ++p) row.push_back({R.idx[p], (int32_t)sp_terms[q], R.val[p]});
        }
        std::stable_sort(row.begin(), row.end(), [](const Contrib &a, const Contrib &b) { return a.k < b.k; });
        for (size_t q = 0; q < row.size();) {
            int64_t k = row[q].k;
            cd v = 0;
            for (; q < row.size() && row[q].k == k; ++q) {
                v += cf[(size_t)row[q].t] * row[q].b;
                MC->cterm.push_back(row[q].t);
                MC->cbase.push_back(row[q].b);
            }
            S->csr.idx.push_back((int32_t)k);
            S->csr.val.push_back(v);
            MC->cptr.push_back((int64_t)MC->cterm.size());
        }
        S->csr.ptr[i + 1] = (int32_t)S->csr.idx.size();
    }
    S->csc = transpose_rows(S->csr, false);
    // CSC entry -> CSR entry (the transpose visits rows in order)
    {
        MC->csc2csr.resize(S->csr.idx.size());
        std::vector<int32_t> pos(S->csc.ptr.begin(), S->csc.ptr.end() - 1);
        for (int64_t r = 0; r < m; ++r)
            for (int32_t p = S->csr.ptr[r]; p < S->csr.ptr[r + 1]; ++p) MC->csc2csr[pos[S->csr.idx[p]]++] = p;
    }
    MC->dcoef = &op.dcoef.p;
    MC->dcoef_on = &op.dcoef_on;
    S->mc = MC;
    op.sums.emplace(key, S);
    out = S;
    return true;
}

// LazySum (src/operators_lazysum.jl:189-238).  Fused mode groups the terms:
// on a density matrix with host-held coefficients all pure-sparse
// LazyTensors and SparseOperators become one summed sparse matrix;
// otherwise the pure-sparse LazyTensors become one gather launch and the
// SparseOperators one summed sparse matrix.  The remaining terms follow with
// beta = 1.
static qob_status apply_ls(const Apply &ap, OpNode &op, const void *in, void *out, cd alpha, cd beta) {
    const int64_t dout = op_out_dim(op, ap.trans);
    if (op.terms.empty() || is_zero(alpha)) return apply_scale(ap, dout, out, beta);
    if (ap.ctx->mode == QOB_MODE_PLAIN) {
        if (op.dcoef_on)
            return fail(QOB_EUNSUPPORTED, "LazySum with device-held coefficients: use the fused mode");
        for (size_t t = 0; t < op.terms.size(); ++t)
            QOB_TRY(apply_op(ap, *op.terms[t], in, out, alpha * op.coef[t], t == 0 ? beta : cd(1, 0)));
        return QOB_OK;
    }
    {
        bool done = false;
        QOB_TRY(apply_qubit(ap, op, in, out, alpha, beta, done));
        if (done) return QOB_OK;
    }
    std::vector<OpNode *> lts;
    std::vector<cd> lcoef;
    std::vector<size_t> lt_terms, sp_terms, rest;
    for (size_t t = 0; t < op.terms.size(); ++t) {
        OpNode &x = *op.terms[t];
        if (lt_pure_sparse(x)) {
            lts.push_back(&x);
            lcoef.push_back(op.coef[t]);
            lt_terms.push_back(t);
        } else if (x.kind == K_SPARSE) {
            sp_terms.push_back(t);
        } else {
            rest.push_back(t);
        }
    }
    cd b = beta;
    const int64_t nvec = ap.lead == 1 ? ap.trail : ap.lead;
    if (!lts.empty() && g_ls_materialize && nvec >= 8 && op.m < (1ll << 31) && op.n < (1ll << 31)) {
        // with device-held coefficients the tables take their values from
        // them (refreshed on the device whenever they change)
        std::shared_ptr<SparseData> SD;
        if (ls_materialize(op, lts, lcoef, lt_terms, sp_terms, SD)) {
            OpNode tmp;
            tmp.kind = K_SPARSE;
            tmp.m = op.m;
            tmp.n = op.n;
            tmp.sp = SD;
            QOB_TRY(apply_sparse(ap, tmp, in, out, alpha, b));
            b = cd(1, 0);
            lts.clear();
            sp_terms.clear();
        }
    }
    if (!lts.empty()) {
        GatherPlanDev P;
        QOB_TRY(lazy_plan(op, "ls" + std::to_string(op.version), lts, lcoef, ap.trans, ap.dt, P,
                          lts.size() == op.terms.size(), ap.s));
        QOB_TRY(apply_gather(ap, P, op_in_dim(op, ap.trans), dout, in, out, alpha, b));
        b = cd(1, 0);
    }
    if (!sp_terms.empty()) {
        // summed sparse matrix sum_t c_t op_t(M_t), built once per
        // coefficient set and cached on the LazySum node
        std::shared_ptr<SparseData> SD;
        {
            std::lock_guard<std::mutex> g(op.mu);
            const std::string k = "sps" + std::to_string(op.version);
            auto it = op.sums.find(k);
            if (it == op.sums.end()) {
                std::vector<std::map<int32_t, cd>> acc((size_t)op.m);
                for (size_t t : sp_terms) {
                    OpNode &x = *op.terms[t];
                    RowLists R = x.sp->rows(x.var);
                    for (int64_t i = 0; i < R.nr; ++i)
                        for (int32_t p = R.ptr[i]; p < R.ptr[i + 1]; ++p) acc[i][R.idx[p]] += op.coef[t] * R.val[p];
                }
                auto S = std::make_shared<SparseData>();
                S->m = op.m;
                S->n = op.n;
                S->csr.nr = op.m;
                S->csr.nc = op.n;
                S->csr.ptr.assign(op.m + 1, 0);
                for (int64_t i = 0; i < op.m; ++i) {
                    for (auto &kv : acc[i]) {
                        S->csr.idx.push_back(kv.first);
                        S->csr.val.push_back(kv.second);
                    }
                    S->csr.ptr[i + 1] = (int32_t)S->csr.idx.size();
                }
                S->csc = transpose_rows(S->csr, false);
                it = op.sums.emplace(k, S).first;
            }
            SD = it->second;
        }
        OpNode tmp;
        tmp.kind = K_SPARSE;
        tmp.m = op.m;
        tmp.n = op.n;
        tmp.sp = SD;
        QOB_TRY(apply_sparse(ap, tmp, in, out, alpha, b));
        b = cd(1, 0);
    }
    for (size_t t : rest) {
        QOB_TRY(apply_op(ap, *op.terms[t], in, out, alpha * op.coef[t], b));
        b = cd(1, 0);
    }
    return QOB_OK;
}

// LazyProduct (src/operators_lazyproduct.jl:103-163): the factors applied in
// turn through two temporaries of the view (ket_l / bra_r / tmp1, tmp2 of the
// reference), the product's factor on the first application and (alpha,
// beta) on the last.  A Ket or a left application runs the last factor
// first; a Bra or a right application (transposed) the first factor first.
static qob_status apply_lp(const Apply &ap, OpNode &op, const void *in, void *out, cd alpha, cd beta) {
    const int64_t dout = op_out_dim(op, ap.trans);
    if (is_zero(alpha)) return apply_scale(ap, dout, out, beta);
    const size_t n = op.terms.size();
    if (n == 1) return apply_op(ap, *op.terms[0], in, out, alpha * op.factor, beta);
    if (ap.depth >= LP_MAX_DEPTH) return fail(QOB_EUNSUPPORTED, "LazyProduct nested more than %d deep", LP_MAX_DEPTH);
    Apply sub = ap;
    sub.depth = ap.depth + 1;
    auto fac = [&](size_t k) -> OpNode & { return *op.terms[ap.trans ? k : n - 1 - k]; };
    const void *src = in;
    for (size_t k = 0; k + 1 < n; ++k) {
        OpNode &f = fac(k);
        const int64_t d = op_out_dim(f, ap.trans);
        void *dst;
        QOB_TRY(ws_get(ap.ctx, ap.s, WS_LP0 + 2 * ap.depth + (int)(k & 1),
                       (size_t)(d * ap.lead * ap.trail) * esize(ap.dt), dst));
        QOB_TRY(apply_op(sub, f, src, dst, k == 0 ? op.factor : cd(1, 0), cd(0, 0)));
        src = dst;
    }
    return apply_op(sub, fac(n - 1), src, out, alpha, beta);
}

static qob_status apply_op(const Apply &ap, OpNode &op, const void *in, void *out, cd alpha, cd beta) {
    switch (op.kind) {
    case K_LP: return apply_lp(ap, op, in, out, alpha, beta);
    case K_SPARSE: return apply_sparse(ap, op, in, out, alpha, beta);
    case K_DENSE: return apply_dense(ap, op, in, out, alpha, beta);
    case K_EYE: return apply_embed(ap, op_out_dim(op, ap.trans), op_in_dim(op, ap.trans), in, out, alpha, beta);
    case K_LT: return apply_lt(ap, op, in, out, alpha, beta);
    case K_LS: return apply_ls(ap, op, in, out, alpha, beta);
    }
    return fail(QOB_EUNSUPPORTED, "unknown operator kind");
}

}  // namespace qob

using namespace qob;

// ====================================================================== ABI
extern "C" {

const char *qob_last_error(void) { return g_err.c_str(); }
const char *qob_version(void) { return "qob_hip 0.1.0 (gfx950)"; }

qob_status qob_ctx_create(int device, qob_ctx *out) {
    if (!out) return fail(QOB_EARG, "null output pointer");
    int n = 0;
    QOB_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(QOB_EARG, "device %d not present (%d devices)", device, n);
    QOB_HIP(hipSetDevice(device));
    QOB_TRY(noalloc_init(device));
    auto *c = new qob_ctx_s();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(QOB_EHIP, "hipStreamCreateWithFlags failed");
    }
    *out = c;
    return QOB_OK;
}

qob_status qob_ctx_destroy(qob_ctx ctx) {
    if (!ctx) return QOB_OK;
    (void)hipDeviceSynchronize();
    {
        std::lock_guard<std::mutex> g(ctx->rmu);
        for (Retired &R : ctx->retired)
            for (hipEvent_t e : R.evs) (void)hipEventDestroy(e);
        ctx->retired.clear();
    }
    ctx->ws.clear();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return QOB_OK;
}

qob_status qob_ctx_set_mode(qob_ctx ctx, qob_mode mode) {
    if (!ctx) return fail(QOB_EARG, "null context");
    if (mode != QOB_MODE_FUSED && mode != QOB_MODE_PLAIN) return fail(QOB_EARG, "bad mode %d", (int)mode);
    ctx->mode = mode;
    return QOB_OK;
}

qob_status qob_ctx_get_stream(qob_ctx ctx, qob_stream *out) {
    if (!ctx || !out) return fail(QOB_EARG, "null argument");
    *out = (qob_stream)ctx->stream;
    return QOB_OK;
}

qob_status qob_dmalloc(qob_ctx ctx, size_t bytes, void **dptr) {
    if (!ctx || !dptr) return fail(QOB_EARG, "null argument");
    if (hipMalloc(dptr, bytes ? bytes : 16) != hipSuccess) {
        *dptr = nullptr;
        (void)hipGetLastError();
        return fail(QOB_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    return QOB_OK;
}

qob_status qob_dfree(qob_ctx ctx, void *dptr) {
    (void)ctx;
    if (dptr) QOB_HIP(hipFree(dptr));
    return QOB_OK;
}

qob_status qob_memcpy_htod(qob_ctx ctx, void *dst, const void *src, size_t bytes, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    hipStream_t st = strm(ctx, s);
    QOB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    QOB_HIP(hipStreamSynchronize(st));
    return QOB_OK;
}

qob_status qob_memcpy_dtoh(qob_ctx ctx, void *dst, const void *src, size_t bytes, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    hipStream_t st = strm(ctx, s);
    QOB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    QOB_HIP(hipStreamSynchronize(st));
    return QOB_OK;
}

qob_status qob_memcpy_dtod(qob_ctx ctx, void *dst, const void *src, size_t bytes, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    QOB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, strm(ctx, s)));
    return QOB_OK;
}

qob_status qob_stream_create(qob_ctx ctx, qob_stream *out) {
    if (!ctx || !out) return fail(QOB_EARG, "null argument");
    hipStream_t s;
    QOB_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->lib_streams.insert(s);
    }
    *out = (qob_stream)s;
    return QOB_OK;
}

qob_status qob_stream_destroy(qob_ctx ctx, qob_stream s) {
    if (!ctx || !s) return fail(QOB_EARG, "null argument");
    QOB_HIP(hipStreamSynchronize((hipStream_t)s));
    std::unique_ptr<Workspace> w;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->ws.find((hipStream_t)s);
        if (it != ctx->ws.end()) {
            w = std::move(it->second);
            ctx->ws.erase(it);
        }
        ctx->lib_streams.erase((hipStream_t)s);
    }
    if (w) {  // a live captured graph may still use this stream's workspace
        std::vector<std::shared_ptr<void>> keep;
        for (auto &b : w->slot) keep.push_back(b.release());
        std::lock_guard<std::mutex> g(ctx->rmu);
        if (ctx->live_graphs > 0) {
            Retired R;
            R.keep = std::move(keep);
            ctx->retired.push_back(std::move(R));
        }
    }
    QOB_HIP(hipStreamDestroy((hipStream_t)s));
    return QOB_OK;
}

qob_status qob_stream_synchronize(qob_ctx ctx, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    QOB_HIP(hipStreamSynchronize(strm(ctx, s)));
    return QOB_OK;
}

// ---------------------------------------------------------------- events
qob_status qob_event_create(qob_ctx ctx, qob_event *out) {
    if (!ctx || !out) return fail(QOB_EARG, "null argument");
    QOB_HIP(hipSetDevice(ctx->device));
    hipEvent_t e = nullptr;
    QOB_HIP(hipEventCreate(&e));
    *out = (qob_event)e;
    return QOB_OK;
}

qob_status qob_event_record(qob_ctx ctx, qob_event ev, qob_stream s) {
    if (!ctx || !ev) return fail(QOB_EARG, "null argument");
    QOB_HIP(hipEventRecord((hipEvent_t)ev, strm(ctx, s)));
    return QOB_OK;
}

qob_status qob_event_query(qob_event ev, int *done) {
    if (!ev || !done) return fail(QOB_EARG, "null argument");
    const hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) {
        *done = 1;
        return QOB_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return QOB_OK;
    }
    return fail(QOB_EHIP, "hipEventQuery: %s", hipGetErrorString(e));
}

qob_status qob_event_synchronize(qob_event ev) {
    if (!ev) return fail(QOB_EARG, "null event");
    QOB_HIP(hipEventSynchronize((hipEvent_t)ev));
    return QOB_OK;
}

qob_status qob_event_elapsed(qob_event start, qob_event stop, float *ms) {
    if (!start || !stop || !ms) return fail(QOB_EARG, "null argument");
    QOB_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return QOB_OK;
}

qob_status qob_event_destroy(qob_event ev) {
    if (ev) (void)hipEventDestroy((hipEvent_t)ev);
    return QOB_OK;
}

// ------------------------------------------------------------- operators
static qob_op wrap(qob_ctx ctx, std::shared_ptr<OpNode> n) {
    auto *h = new qob_op_s();
    h->node = std::move(n);
    h->ctx = ctx;
    return h;
}

qob_status qob_sparse_create(qob_ctx ctx, int64_t m, int64_t n, const int64_t *colptr1, const int64_t *rowval1,
                             const qob_c64 *nzval, qob_op *out) {
    if (!ctx || !out || !colptr1) return fail(QOB_EARG, "null argument");
    if (m < 0 || n < 0 || m >= (1ll << 31) || n >= (1ll << 31)) return fail(QOB_EARG, "bad dims %lldx%lld", (long long)m, (long long)n);
    const int64_t nnz = colptr1[n] - 1;
    if (colptr1[0] != 1 || nnz < 0) return fail(QOB_EARG, "colptr must be 1-based");
    if (nnz > 0 && (!rowval1 || !nzval)) return fail(QOB_EARG, "null rowval/nzval");
    auto S = std::make_shared<SparseData>();
    S->m = m;
    S->n = n;
    S->csc.nr = n;
    S->csc.nc = m;
    S->csc.ptr.resize(n + 1);
    for (int64_t j = 0; j <= n; ++j) {
        S->csc.ptr[j] = (int32_t)(colptr1[j] - 1);
        if (j > 0 && colptr1[j] < colptr1[j - 1]) return fail(QOB_EARG, "colptr not monotone");
    }
    S->csc.idx.resize(nnz);
    S->csc.val.resize(nnz);
    for (int64_t p = 0; p < nnz; ++p) {
        const int64_t r = rowval1[p] - 1;
        if (r < 0 || r >= m) return fail(QOB_EARG, "rowval[%lld] = %lld out of range", (long long)p, (long long)rowval1[p]);
        S->csc.idx[p] = (int32_t)r;
        S->csc.val[p] = cdv(nzval[p]);
    }
    S->csr = transpose_rows(S->csc, false);
    auto node = std::make_shared<OpNode>();
    node->kind = K_SPARSE;
    node->m = m;
    node->n = n;
    node->sp = S;
    *out = wrap(ctx, node);
    return QOB_OK;
}

qob_status qob_dense_create(qob_ctx ctx, int64_t m, int64_t n, const qob_c64 *data, qob_op *out) {
    if (!ctx || !out) return fail(QOB_EARG, "null argument");
    if (m < 0 || n < 0) return fail(QOB_EARG, "bad dims");
    if (m * n > 0 && !data) return fail(QOB_EARG, "null data");
    auto D = std::make_shared<DenseData>();
    D->m = m;
    D->n = n;
    D->a.resize((size_t)(m * n));
    for (int64_t i = 0; i < m * n; ++i) D->a[i] = cdv(data[i]);
    auto node = std::make_shared<OpNode>();
    node->kind = K_DENSE;
    node->m = m;
    node->n = n;
    node->de = D;
    *out = wrap(ctx, node);
    return QOB_OK;
}

qob_status qob_eye_create(qob_ctx ctx, int64_t m, int64_t n, qob_op *out) {
    if (!ctx || !out) return fail(QOB_EARG, "null argument");
    if (m < 0 || n < 0) return fail(QOB_EARG, "bad dims");
    auto node = std::make_shared<OpNode>();
    node->kind = K_EYE;
    node->m = m;
    node->n = n;
    *out = wrap(ctx, node);
    return QOB_OK;
}

static std::shared_ptr<OpNode> adjoint_node(const std::shared_ptr<OpNode> &a) {
    auto r = std::make_shared<OpNode>();
    r->kind = a->kind;
    r->m = a->n;
    r->n = a->m;
    switch (a->kind) {
    case K_SPARSE: r->sp = a->sp; r->var = a->var ^ V_H; break;
    case K_DENSE: r->de = a->de; r->var = a->var ^ V_H; break;
    case K_EYE: break;
    case K_LT:
        r->shl = a->shr;
        r->shr = a->shl;
        r->factor = std::conj(a->factor);
        r->fsite = a->fsite;
        for (auto &f : a->fop) r->fop.push_back(adjoint_node(f));
        break;
    case K_LS: {
        for (auto &c : a->coef) r->coef.push_back(std::conj(c));
        for (auto &t : a->terms) r->terms.push_back(adjoint_node(t));
        std::vector<double2> dc(r->coef.size());
        for (size_t t = 0; t < dc.size(); ++t) dc[t] = d2(r->coef[t]);
        if (!dc.empty()) (void)r->dcoef.upload(dc.data(), dc.size() * sizeof(double2));
        break;
    }
    case K_LP:  // (A1 ... An)^dag = An^dag ... A1^dag
        r->factor = std::conj(a->factor);
        for (auto it = a->terms.rbegin(); it != a->terms.rend(); ++it) r->terms.push_back(adjoint_node(*it));
        break;
    }
    return r;
}

qob_status qob_lazyproduct_create(qob_ctx ctx, int nops, const qob_op *operators, qob_c64 factor, qob_op *out) {
    if (!ctx || !out || !operators) return fail(QOB_EARG, "null argument");
    if (nops < 1) return fail(QOB_EARG, "LazyProduct needs at least one operator!");
    auto node = std::make_shared<OpNode>();
    node->kind = K_LP;
    node->factor = cdv(factor);
    for (int i = 0; i < nops; ++i) {
        if (!operators[i]) return fail(QOB_EARG, "null operator %d", i + 1);
        auto &x = operators[i]->node;
        if (i > 0 && node->terms.back()->n != x->m)
            return fail(QOB_EBASIS, "LazyProduct: operator %d is %lldx%lld after one with %lld columns", i + 1,
                        (long long)x->m, (long long)x->n, (long long)node->terms.back()->n);
        node->terms.push_back(x);
    }
    node->m = node->terms.front()->m;
    node->n = node->terms.back()->n;
    *out = wrap(ctx, node);
    return QOB_OK;
}

qob_status qob_op_adjoint(qob_ctx ctx, qob_op a, qob_op *out) {
    if (!ctx || !a || !out) return fail(QOB_EARG, "null argument");
    *out = wrap(ctx, adjoint_node(a->node));
    return QOB_OK;
}

qob_status qob_lazytensor_create(qob_ctx ctx, int nsites, const int64_t *shape_l, const int64_t *shape_r,
                                 int nfactors, const int64_t *indices1, const qob_op *factors, qob_c64 factor,
                                 qob_op *out) {
    if (!ctx || !out || nsites < 1 || !shape_l || !shape_r) return fail(QOB_EARG, "bad arguments");
    if (nfactors < 0 || (nfactors > 0 && (!indices1 || !factors))) return fail(QOB_EARG, "bad factor arguments");
    auto node = std::make_shared<OpNode>();
    node->kind = K_LT;
    node->shl.assign(shape_l, shape_l + nsites);
    node->shr.assign(shape_r, shape_r + nsites);
    node->m = prod(node->shl, 0, nsites);
    node->n = prod(node->shr, 0, nsites);
    for (int s = 0; s < nsites; ++s)
        if (shape_l[s] < 1 || shape_r[s] < 1) return fail(QOB_EARG, "site dims must be positive");
    node->factor = cdv(factor);
    for (int i = 0; i < nfactors; ++i) {
        const int64_t s = indices1[i] - 1;
        if (s < 0 || s >= nsites) return fail(QOB_EARG, "index %lld out of range 1..%d", (long long)indices1[i], nsites);
        if (i > 0 && indices1[i] <= indices1[i - 1])
            return fail(QOB_EARG, "indices must be sorted and unique");  // src/operators_lazytensor.jl:26
        if (!factors[i]) return fail(QOB_EARG, "null factor");
        const OpNode &f = *factors[i]->node;
        if (f.kind == K_LT || f.kind == K_LS || f.kind == K_LP)
            return fail(QOB_EUNSUPPORTED, "LazyTensor factors must be data operators");
        if (f.m != shape_l[s] || f.n != shape_r[s])
            return fail(QOB_EBASIS, "factor %d is %lldx%lld but site %lld is %lldx%lld", i + 1, (long long)f.m,
                        (long long)f.n, (long long)(s + 1), (long long)shape_l[s], (long long)shape_r[s]);
        node->fsite.push_back((int)s);
        node->fop.push_back(factors[i]->node);
    }
    *out = wrap(ctx, node);
    return QOB_OK;
}

qob_status qob_lazysum_create(qob_ctx ctx, int64_t m, int64_t n, int nterms, const qob_c64 *factors,
                              const qob_op *operators, qob_op *out) {
    if (!ctx || !out || nterms < 0) return fail(QOB_EARG, "bad arguments");
    if (nterms > 0 && (!factors || !operators)) return fail(QOB_EARG, "null terms");
    auto node = std::make_shared<OpNode>();
    node->kind = K_LS;
    node->m = m;
    node->n = n;
    const OpNode *first_lt = nullptr;
    for (int t = 0; t < nterms; ++t) {
        if (!operators[t]) return fail(QOB_EARG, "null term");
        auto &x = operators[t]->node;
        if (x->m != m || x->n != n) return fail(QOB_EBASIS, "term %d is %lldx%lld, sum is %lldx%lld", t + 1,
                                                (long long)x->m, (long long)x->n, (long long)m, (long long)n);
        if (x->kind == K_LT) {
            if (first_lt && (first_lt->shl != x->shl || first_lt->shr != x->shr))
                return fail(QOB_EBASIS, "LazySum terms on different composite bases");
            first_lt = x.get();
        }
        node->coef.push_back(cdv(factors[t]));
        node->terms.push_back(x);
    }
    {
        std::vector<double2> dc(node->coef.size());
        for (size_t t = 0; t < dc.size(); ++t) dc[t] = d2(node->coef[t]);
        if (dc.empty()) dc.push_back(make_double2(0, 0));
        QOB_TRY(node->dcoef.upload(dc.data(), dc.size() * sizeof(double2)));
    }
    *out = wrap(ctx, node);
    return QOB_OK;
}

qob_status qob_lazysum_set_factors(qob_ctx ctx, qob_op sum, const qob_c64 *factors) {
    if (!ctx || !sum || sum->node->kind != K_LS) return fail(QOB_EARG, "not a LazySum");
    OpNode &n = *sum->node;
    if (!n.terms.empty() && !factors) return fail(QOB_EARG, "null factors");
    // plans of the old coefficients may still be read by queued work or by a
    // captured graph: they are retired (freed once that work is done and no
    // graph is alive); the next application builds plans of the new values
    std::vector<std::shared_ptr<void>> old;
    {
        std::lock_guard<std::mutex> g(n.mu);
        for (size_t t = 0; t < n.terms.size(); ++t) n.coef[t] = cdv(factors[t]);
        n.version++;
        for (auto &kv : n.gplans) old.push_back(kv.second);
        for (auto &kv : n.sums) old.push_back(kv.second);
        for (auto &kv : n.qplans) old.push_back(kv.second);
        n.gplans.clear();
        n.sums.clear();
        n.qplans.clear();
        // the device copy of the coefficients is read only while dcoef_on
        // (refresh kernels); qob_lazysum_set_factors_d rewrites it
        n.dcoef_on = false;
    }
    retire(ctx, std::move(old));
    return QOB_OK;
}

qob_status qob_lazysum_set_factors_d(qob_ctx ctx, qob_op sum, const void *dfactors, qob_stream s) {
    if (!ctx || !sum || sum->node->kind != K_LS) return fail(QOB_EARG, "not a LazySum");
    OpNode &n = *sum->node;
    if (n.terms.empty()) return QOB_OK;
    if (!dfactors) return fail(QOB_EARG, "null factors");
    for (auto &t : n.terms)
        if (!lt_pure_sparse(*t))
            return fail(QOB_EUNSUPPORTED, "device-held coefficients need every term to be an all-sparse LazyTensor");
    const hipStream_t st = strm(ctx, s);
    std::lock_guard<std::mutex> g(n.mu);
    if (!n.dcoef.p) QOB_TRY(n.dcoef.alloc(n.terms.size() * sizeof(double2)));
    hipError_t e = launch_copy(QOB_C64, (int64_t)n.terms.size(), dfactors, n.dcoef.p, st);
    if (e != hipSuccess) return fail(QOB_EHIP, "coefficient copy: %s", hipGetErrorString(e));
    for (auto &kv : n.qplans) {
        QubitPlan &P = *kv.second;
        if (!P.ok) continue;
        if (!P.rmap) return fail(QOB_EUNSUPPORTED, "qubit plan without a coefficient map");
        QOB_TRY(run_refresh(*P.rmap, P.dt, n.dcoef.p, P.vals.p, P.prog.p, st));
    }
    for (auto &kv : n.gplans) {
        GatherPlan &P = *kv.second;
        if (!P.rmap) return fail(QOB_EUNSUPPORTED, "gather plan without a coefficient map");
        QOB_TRY(run_refresh(*P.rmap, P.dt, n.dcoef.p, P.vals.p, nullptr, st));
    }
    n.dcoef_on = true;
    for (auto &kv : n.sums)
        if (kv.second && kv.second->mc) QOB_TRY(refresh_mat(*kv.second, st));
    return QOB_OK;
}

qob_status qob_op_shape(qob_op a, int64_t *m, int64_t *n) {
    if (!a) return fail(QOB_EARG, "null operator");
    if (m) *m = a->node->m;
    if (n) *n = a->node->n;
    return QOB_OK;
}

qob_status qob_op_release(qob_op a) {
    if (!a) return QOB_OK;
    // a captured graph may still use the operator's plans and tables: while
    // any graph of its context is alive the node is retired, not destroyed
    if (a->ctx) {
        bool graphs;
        {
            std::lock_guard<std::mutex> g(a->ctx->rmu);
            graphs = a->ctx->live_graphs > 0;
        }
        if (graphs) retire(a->ctx, {std::shared_ptr<void>(a->node)});
    }
    delete a;
    return QOB_OK;
}

// ------------------------------------------------------------------ mul!
static qob_status check_common(qob_ctx ctx, qob_dtype dt, qob_op a, const void *b, void *r) {
    if (!ctx) return fail(QOB_EARG, "null context");
    if (!a) return fail(QOB_EARG, "null operator");
    if (dt != QOB_C64 && dt != QOB_C32) return fail(QOB_EARG, "bad dtype %d", (int)dt);
    if (b == r && b) return fail(QOB_EALIAS, "output must not be aliased with input");
    return QOB_OK;
}

qob_status qob_mul_ket(qob_ctx ctx, qob_dtype dt, qob_op a, const void *b, int64_t nb, void *result, int64_t nr,
                       qob_c64 alpha, qob_c64 beta, qob_stream s) {
    QOB_TRY(check_common(ctx, dt, a, b, result));
    OpNode &op = *a->node;
    if (op.n != nb || op.m != nr)
        return fail(QOB_EDIM, "Output dimensions do not match A*B: op %lldx%lld, b %lld, result %lld",
                    (long long)op.m, (long long)op.n, (long long)nb, (long long)nr);
    Apply ap{ctx, strm(ctx, s), dt, 1, 1, false};
    return apply_op(ap, op, b, result, cdv(alpha), cdv(beta));
}

qob_status qob_mul_bra(qob_ctx ctx, qob_dtype dt, const void *a, int64_t na, qob_op b, void *result, int64_t nr,
                       qob_c64 alpha, qob_c64 beta, qob_stream s) {
    QOB_TRY(check_common(ctx, dt, b, a, result));
    OpNode &op = *b->node;
    if (op.m != na || op.n != nr)
        return fail(QOB_EDIM, "Output dimensions do not match a*B: a %lld, op %lldx%lld, result %lld", (long long)na,
                    (long long)op.m, (long long)op.n, (long long)nr);
    Apply ap{ctx, strm(ctx, s), dt, 1, 1, true};
    return apply_op(ap, op, a, result, cdv(alpha), cdv(beta));
}

qob_status qob_mul_op_left(qob_ctx ctx, qob_dtype dt, qob_op a, const void *B, int64_t rows_b, int64_t cols, void *R,
                           int64_t rows_r, qob_c64 alpha, qob_c64 beta, qob_stream s) {
    QOB_TRY(check_common(ctx, dt, a, B, R));
    OpNode &op = *a->node;
    if (op.n != rows_b || op.m != rows_r || cols < 0)
        return fail(QOB_EDIM, "Output dimensions do not match A*B: op %lldx%lld, B %lldx%lld, R %lldx%lld",
                    (long long)op.m, (long long)op.n, (long long)rows_b, (long long)cols, (long long)rows_r,
                    (long long)cols);
    if (cols == 0) return QOB_OK;
    Apply ap{ctx, strm(ctx, s), dt, 1, cols, false};
    return apply_op(ap, op, B, R, cdv(alpha), cdv(beta));
}

qob_status qob_mul_op_right(qob_ctx ctx, qob_dtype dt, const void *B, int64_t rows, int64_t cols_b, qob_op a, void *R,
                            int64_t cols_r, qob_c64 alpha, qob_c64 beta, qob_stream s) {
    QOB_TRY(check_common(ctx, dt, a, B, R));
    OpNode &op = *a->node;
    if (op.m != cols_b || op.n != cols_r || rows < 0)
        return fail(QOB_EDIM, "Output dimensions do not match B*A: B %lldx%lld, op %lldx%lld, R %lldx%lld",
                    (long long)rows, (long long)cols_b, (long long)op.m, (long long)op.n, (long long)rows,
                    (long long)cols_r);
    if (rows == 0) return QOB_OK;
    Apply ap{ctx, strm(ctx, s), dt, rows, 1, true};
    if (rows == 1) ap.lead = 1;  // a single row is a Bra: contiguous gather axis
    return apply_op(ap, op, B, R, cdv(alpha), cdv(beta));
}

// State and result of different element types: the product runs in
// ComplexF64 (the promoted type, src/operators_lazytensor.jl:32-33), the
// ComplexF32 operand widened into a workspace first (exactly) and a
// ComplexF32 result rounded once at the end.
}  // extern "C"

template <typename F>
static qob_status run_mixed(qob_ctx ctx, qob_stream s, qob_dtype dt_b, const void *b, int64_t nb, qob_dtype dt_r,
                            void *r, int64_t nr, qob_c64 beta, F &&call) {
    if ((dt_b != QOB_C64 && dt_b != QOB_C32) || (dt_r != QOB_C64 && dt_r != QOB_C32))
        return fail(QOB_EARG, "bad dtype %d / %d", (int)dt_b, (int)dt_r);
    if (dt_b == dt_r) return call(dt_r, b, r);
    if (b == r && b) return fail(QOB_EALIAS, "output must not be aliased with input");
    hipStream_t st = strm(ctx, s);
    const void *bw = b;
    void *rw = r;
    if (dt_b == QOB_C32 && nb > 0) {
        void *t;
        QOB_TRY(ws_get(ctx, st, WS_MIXB, (size_t)nb * 16, t));
        QOB_HIP(launch_convert(QOB_C32, b, QOB_C64, t, nb, st));
        bw = t;
    }
    if (dt_r == QOB_C32 && nr > 0) {
        void *t;
        QOB_TRY(ws_get(ctx, st, WS_MIXR, (size_t)nr * 16, t));
        if (!is_zero(cdv(beta))) QOB_HIP(launch_convert(QOB_C32, r, QOB_C64, t, nr, st));
        rw = t;
    }
    QOB_TRY(call(QOB_C64, bw, rw));
    if (dt_r == QOB_C32 && nr > 0) QOB_HIP(launch_convert(QOB_C64, rw, QOB_C32, r, nr, st));
    return QOB_OK;
}

extern "C" {

qob_status qob_mul_ket_x(qob_ctx ctx, qob_dtype dt_b, qob_op a, const void *b, int64_t nb, qob_dtype dt_r,
                         void *result, int64_t nr, qob_c64 alpha, qob_c64 beta, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    return run_mixed(ctx, s, dt_b, b, nb, dt_r, result, nr, beta, [&](qob_dtype dt, const void *bb, void *rr) {
        return qob_mul_ket(ctx, dt, a, bb, nb, rr, nr, alpha, beta, s);
    });
}

qob_status qob_mul_bra_x(qob_ctx ctx, qob_dtype dt_a, const void *a, int64_t na, qob_op b, qob_dtype dt_r,
                         void *result, int64_t nr, qob_c64 alpha, qob_c64 beta, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    return run_mixed(ctx, s, dt_a, a, na, dt_r, result, nr, beta, [&](qob_dtype dt, const void *aa, void *rr) {
        return qob_mul_bra(ctx, dt, aa, na, b, rr, nr, alpha, beta, s);
    });
}

qob_status qob_mul_op_left_x(qob_ctx ctx, qob_dtype dt_b, qob_op a, const void *B, int64_t rows_b, int64_t cols,
                             qob_dtype dt_r, void *R, int64_t rows_r, qob_c64 alpha, qob_c64 beta, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    if (rows_b < 0 || rows_r < 0 || cols < 0) return fail(QOB_EDIM, "negative dimension");
    return run_mixed(ctx, s, dt_b, B, rows_b * cols, dt_r, R, rows_r * cols, beta,
                     [&](qob_dtype dt, const void *bb, void *rr) {
                         return qob_mul_op_left(ctx, dt, a, bb, rows_b, cols, rr, rows_r, alpha, beta, s);
                     });
}

qob_status qob_mul_op_right_x(qob_ctx ctx, qob_dtype dt_b, const void *B, int64_t rows, int64_t cols_b, qob_op a,
                              qob_dtype dt_r, void *R, int64_t cols_r, qob_c64 alpha, qob_c64 beta, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    if (rows < 0 || cols_b < 0 || cols_r < 0) return fail(QOB_EDIM, "negative dimension");
    return run_mixed(ctx, s, dt_b, B, rows * cols_b, dt_r, R, rows * cols_r, beta,
                     [&](qob_dtype dt, const void *bb, void *rr) {
                         return qob_mul_op_right(ctx, dt, bb, rows, cols_b, a, rr, cols_r, alpha, beta, s);
                     });
}

qob_status qob_gemm(qob_ctx ctx, qob_dtype dt, int transa, int64_t m, int64_t n, int64_t k, qob_c64 alpha,
                    const void *A, const void *B, qob_c64 beta, void *C, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    if (C == A || C == B) return fail(QOB_EALIAS, "output must not be aliased with input");
    if (m < 0 || n < 0 || k < 0) return fail(QOB_EDIM, "negative dimension");
    hipStream_t st = strm(ctx, s);
    hipError_t e;
    if (k == 0) {  // no product: C = beta C (beta = 0 writes zeros)
        e = launch_scale(dt, m * n, d2(beta), C, st);
    } else if (n == 1 && !transa) {  // matrix-vector (src/operators_dense.jl:395)
        void *ws = nullptr;
        const int64_t ns = zgemv_splits(m, k);
        if (ns > 1) QOB_TRY(ws_get(ctx, st, WS_GEMV, (size_t)(ns * m) * 16, ws));
        e = launch_zgemv(dt, A, m, k, B, C, ws, d2(alpha), d2(beta), st);
    } else if (n == 1) {  // transpose(A) x (the Bra form, src/operators_dense.jl:396)
        e = launch_zgemv_t(dt, A, k, m, B, C, d2(alpha), d2(beta), st);
    } else  // op(A) (m x k): A column-major m x k, or k x m read transposed
        e = launch_zgemm(dt, m, n, k, A, transa ? k : 1, transa ? 1 : m, 0, B, 1, k, 0, C, 1, m, 0, 1, d2(alpha),
                         d2(beta), st);
    if (e != hipSuccess) return fail(QOB_EHIP, "gemm launch: %s", hipGetErrorString(e));
    return QOB_OK;
}

// ----------------------------------------------------- elementwise / reduce
qob_status qob_scale(qob_ctx ctx, qob_dtype dt, int64_t n, qob_c64 a, void *x, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    hipError_t e = launch_scale(dt, n, d2(a), x, strm(ctx, s));
    if (e != hipSuccess) return fail(QOB_EHIP, "scale: %s", hipGetErrorString(e));
    return QOB_OK;
}

qob_status qob_copy(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x, void *y, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    if (x == y) return fail(QOB_EALIAS, "copy with x == y");
    hipError_t e = launch_copy(dt, n, x, y, strm(ctx, s));
    if (e != hipSuccess) return fail(QOB_EHIP, "copy: %s", hipGetErrorString(e));
    return QOB_OK;
}

static int *tune_slot(const char *name) {
    static const struct {
        const char *n;
        int *p;
    } tbl[] = {{"sell_variant", &g_sell_variant}, {"sell_cj", &g_sell_cj},     {"sell_jchunk", &g_sell_jchunk},
               {"csc_rb", &g_csc_rb},             {"csc_cc", &g_csc_cc},       {"csc_remap", &g_csc_remap},
               {"nt_store", &g_nt_store},         {"copy_variant", &g_copy_variant},
               {"fused_lt", &g_fused_lt},         {"qubit_enable", &g_qubit_enable},
               {"qubit_tbits", &g_qubit_tbits},   {"qubit_nth", &g_qubit_nth},
               {"qubit_grid", &g_qubit_grid},     {"qubit_remap", &g_qubit_remap},
               {"qubit_split", &g_qubit_split},   {"mode_kernel", &g_mode_kernel}, {"fused_reg", &g_fused_reg},
               {"qubit_ntfar", &g_qubit_ntfar}, {"qubit_order", &g_qubit_order},
               {"qubit_dbuf", &g_qubit_dbuf}, {"qubit_pair", &g_qubit_pair},
               {"fused_gpw", &g_fused_gpw},       {"csc_rows", &g_csc_rows},
               {"sell_unr", &g_sell_unr},         {"csc_order", &g_csc_order},
               {"sell_align", &g_sell_align},     {"csc_waves", &g_csc_waves},
               {"csc_sweep", &g_csc_sweep},       {"sell_closed", &g_sell_closed}, {"sell_cline", &g_sell_cline}, {"sell_cpair", &g_sell_cpair},
               {"sell_cjchunk", &g_sell_cjchunk}, {"csc_groups", &g_csc_groups},
               {"csc_gkernel", &g_csc_gkernel},   {"csc_gremap", &g_csc_gremap},
               {"sell_cremap", &g_sell_cremap},   {"ls_materialize", &g_ls_materialize},
               {"fused_tiles", &g_fused_tiles},   {"fused_resident", &g_fused_resident}};
    for (auto &e : tbl)
        if (name && strcmp(name, e.n) == 0) return e.p;
    return nullptr;
}

qob_status qob_tune_set(const char *name, int64_t value) {
    int *p = tune_slot(name);
    if (!p) return fail(QOB_EARG, "unknown tunable '%s'", name ? name : "(null)");
    if (value < 0 || value > (1 << 20)) return fail(QOB_EARG, "tunable value %lld out of range", (long long)value);
    if (!strcmp(name, "csc_rb") && value != 64 && value != 128 && value != 256)
        return fail(QOB_EARG, "csc_rb must be 64, 128 or 256");
    if (!strcmp(name, "sell_cj") && value != 1 && value != 2 && value != 4 && value != 8)
        return fail(QOB_EARG, "sell_cj must be 1, 2, 4 or 8");
    if (!strcmp(name, "qubit_tbits") && value != 0 && (value < 10 || value > 13))
        return fail(QOB_EARG, "qubit_tbits must be 0 (by element type) or in 10..13");
    if (!strcmp(name, "mode_kernel") && value > 2)
        return fail(QOB_EARG, "mode_kernel must be 0 (automatic), 1 (tiled GEMM for every dense mode product) or 2 "
                              "(fibre-group kernel for leading dimensions below 16 as well)");
    if (!strcmp(name, "sell_unr") && value != 2 && value != 4 && value != 8)
        return fail(QOB_EARG, "sell_unr must be 2, 4 or 8");
    if (!strcmp(name, "csc_sweep") && value > 2) return fail(QOB_EARG, "csc_sweep must be 0, 1 or 2");
    if (!strcmp(name, "csc_waves") && value != 1 && value != 2 && value != 4)
        return fail(QOB_EARG, "csc_waves must be 1, 2 or 4");
    if (!strcmp(name, "csc_rows") && value != 16 && value != 32 && value != 64)
        return fail(QOB_EARG, "csc_rows must be 16, 32 or 64");
    if (!strcmp(name, "fused_gpw") && value != 0 && value != 1 && value != 2 && value != 4)
        return fail(QOB_EARG, "fused_gpw must be 0 (automatic), 1, 2 or 4");
    if (!strcmp(name, "qubit_split") && value > 1) return fail(QOB_EARG, "qubit_split must be 0 or 1");
    if (!strcmp(name, "qubit_remap") && value > 2) return fail(QOB_EARG, "qubit_remap must be 0, 1 or 2");
    if (!strcmp(name, "qubit_ntfar") && value > 8) return fail(QOB_EARG, "qubit_ntfar must be in 0..8");
    if (!strcmp(name, "qubit_order") && value > 2) return fail(QOB_EARG, "qubit_order must be 0, 1 or 2");
    if ((!strcmp(name, "csc_order") || !strcmp(name, "sell_align") || !strcmp(name, "sell_closed") ||
         !strcmp(name, "sell_cline") || !strcmp(name, "sell_cpair") ||
         !strcmp(name, "csc_groups") || !strcmp(name, "csc_gkernel") || !strcmp(name, "csc_gremap") ||
         !strcmp(name, "sell_cremap") || !strcmp(name, "ls_materialize") || !strcmp(name, "fused_tiles") ||
         !strcmp(name, "fused_resident") || !strcmp(name, "fused_reg") || !strcmp(name, "qubit_dbuf") ||
         !strcmp(name, "qubit_pair")) &&
        value > 1)
        return fail(QOB_EARG, "%s must be 0 or 1", name);
    if (!strcmp(name, "qubit_nth") && value != 0 && value != 256 && value != 512 && value != 1024)
        return fail(QOB_EARG, "qubit_nth must be 0 (by element type), 256, 512 or 1024");
    if ((!strcmp(name, "csc_cc") || !strcmp(name, "sell_jchunk") || !strcmp(name, "sell_cjchunk")) && value < 1)
        return fail(QOB_EARG, "%s must be >= 1", name);
    *p = (int)value;
    return QOB_OK;
}

qob_status qob_csc_column_order(int64_t n, int64_t xrows, const int32_t *colptr, const int32_t *rowind,
                                int32_t *order, int *used) {
    if (n < 0 || xrows < 0 || (n > 0 && (!colptr || !order)) || !used) return fail(QOB_EARG, "null argument");
    if (n >= (1ll << 31) || xrows >= (1ll << 31)) return fail(QOB_EARG, "dimension out of range");
    RowLists X;
    X.nr = n;
    X.nc = xrows;
    X.ptr.assign(colptr, colptr + n + 1);
    if (X.ptr[0] != 0) return fail(QOB_EARG, "colptr[0] must be 0");
    for (int64_t c = 0; c < n; ++c)
        if (X.ptr[c + 1] < X.ptr[c]) return fail(QOB_EARG, "colptr not monotone at %lld", (long long)c);
    if (X.ptr[n] > 0 && !rowind) return fail(QOB_EARG, "null rowind");
    X.idx.assign(rowind, rowind + X.ptr[n]);
    for (int32_t r : X.idx)
        if (r < 0 || r >= xrows) return fail(QOB_EARG, "row index %d out of range", r);
    int64_t bw = -1;
    const std::vector<int32_t> ord = csc_column_order(X, xrows, &bw);
    *used = ord.empty() ? 0 : 1;
    for (int64_t c = 0; c < n; ++c) order[c] = ord.empty() ? (int32_t)c : ord[c];
    return QOB_OK;
}

qob_status qob_tune_get(const char *name, int64_t *value) {
    int *p = tune_slot(name);
    if (!p || !value) return fail(QOB_EARG, "unknown tunable '%s'", name ? name : "(null)");
    *value = *p;
    return QOB_OK;
}

qob_status qob_axpy(qob_ctx ctx, qob_dtype dt, int64_t n, qob_c64 a, const void *x, void *y, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    if (x == y) return fail(QOB_EALIAS, "axpy with x == y");
    hipError_t e = launch_axpy(dt, n, d2(a), x, y, strm(ctx, s));
    if (e != hipSuccess) return fail(QOB_EHIP, "axpy: %s", hipGetErrorString(e));
    return QOB_OK;
}

qob_status qob_convert(qob_ctx ctx, qob_dtype from, const void *src, qob_dtype to, void *dst, int64_t n, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    hipError_t e = launch_convert(from, src, to, dst, n, strm(ctx, s));
    if (e != hipSuccess) return fail(QOB_EHIP, "convert: %s", hipGetErrorString(e));
    return QOB_OK;
}

static qob_status red_ws(qob_ctx ctx, hipStream_t st, int64_t n, double2 *&part) {
    void *p;
    QOB_TRY(ws_get(ctx, st, WS_RED, (size_t)reduce_blocks(n) * sizeof(double2), p));
    part = (double2 *)p;
    return QOB_OK;
}

qob_status qob_dot_d(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x, const void *y, void *dout, qob_stream s) {
    if (!ctx || !dout) return fail(QOB_EARG, "null argument");
    hipStream_t st = strm(ctx, s);
    double2 *part;
    QOB_TRY(red_ws(ctx, st, n, part));
    hipError_t e = launch_dot(dt, n, x, y, part, (double2 *)dout, st);
    if (e != hipSuccess) return fail(QOB_EHIP, "dot: %s", hipGetErrorString(e));
    return QOB_OK;
}

qob_status qob_norm_d(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x, void *dout, qob_stream s) {
    if (!ctx || !dout) return fail(QOB_EARG, "null argument");
    hipStream_t st = strm(ctx, s);
    double2 *part;
    QOB_TRY(red_ws(ctx, st, n, part));
    hipError_t e = launch_nrm2(dt, n, x, part, (double *)dout, st);
    if (e != hipSuccess) return fail(QOB_EHIP, "norm: %s", hipGetErrorString(e));
    return QOB_OK;
}

qob_status qob_trace_d(qob_ctx ctx, qob_dtype dt, int64_t n, const void *A, void *dout, qob_stream s) {
    if (!ctx || !dout) return fail(QOB_EARG, "null argument");
    hipStream_t st = strm(ctx, s);
    double2 *part;
    QOB_TRY(red_ws(ctx, st, n, part));
    hipError_t e = launch_trace(dt, n, A, part, (double2 *)dout, st);
    if (e != hipSuccess) return fail(QOB_EHIP, "trace: %s", hipGetErrorString(e));
    return QOB_OK;
}

static qob_status scal_ws(qob_ctx ctx, hipStream_t st, void *&p) { return ws_get(ctx, st, WS_SCAL, 64, p); }

}  // extern "C"

template <typename F>
static qob_status reduce_to_host(qob_ctx ctx, qob_stream s, void *host, size_t bytes, F f) {
    hipStream_t st = strm(ctx, s);
    void *d;
    QOB_TRY(scal_ws(ctx, st, d));
    QOB_TRY(f(d));
    QOB_HIP(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, st));
    QOB_HIP(hipStreamSynchronize(st));
    return QOB_OK;
}

extern "C" {

qob_status qob_dot(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x, const void *y, qob_c64 *out, qob_stream s) {
    if (!ctx || !out) return fail(QOB_EARG, "null argument");
    return reduce_to_host(ctx, s, out, 16, [&](void *d) { return qob_dot_d(ctx, dt, n, x, y, d, s); });
}

qob_status qob_norm(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x, double *out, qob_stream s) {
    if (!ctx || !out) return fail(QOB_EARG, "null argument");
    return reduce_to_host(ctx, s, out, 8, [&](void *d) { return qob_norm_d(ctx, dt, n, x, d, s); });
}

qob_status qob_trace(qob_ctx ctx, qob_dtype dt, int64_t n, const void *A, qob_c64 *out, qob_stream s) {
    if (!ctx || !out) return fail(QOB_EARG, "null argument");
    return reduce_to_host(ctx, s, out, 16, [&](void *d) { return qob_trace_d(ctx, dt, n, A, d, s); });
}

qob_status qob_normalize_ket(qob_ctx ctx, qob_dtype dt, int64_t n, void *x, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    hipStream_t st = strm(ctx, s);
    void *d;
    QOB_TRY(scal_ws(ctx, st, d));
    QOB_TRY(qob_norm_d(ctx, dt, n, x, d, s));
    hipError_t e = launch_scale_by_inv_norm(dt, n, (const double *)d, x, st);
    if (e != hipSuccess) return fail(QOB_EHIP, "normalize: %s", hipGetErrorString(e));
    return QOB_OK;
}

qob_status qob_normalize_op(qob_ctx ctx, qob_dtype dt, int64_t n, void *A, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    hipStream_t st = strm(ctx, s);
    void *d;
    QOB_TRY(scal_ws(ctx, st, d));
    QOB_TRY(qob_trace_d(ctx, dt, n, A, d, s));
    hipError_t e = launch_scale_by_inv_trace(dt, n, n, (const double2 *)d, A, st);
    if (e != hipSuccess) return fail(QOB_EHIP, "normalize: %s", hipGetErrorString(e));
    return QOB_OK;
}

// sum_J conj(u[J]) (op v)[J] without writing op v, when op has a fused
// single-pass form: a qubit program or a square all-sparse gather plan.
// done = false: no fused form, nothing launched.
static qob_status fused_bilinear(qob_ctx ctx, hipStream_t st, qob_dtype dt, OpNode &op, const void *u,
                                 const void *v, int64_t n, double2 *dout, bool &done) {
    done = false;
    if (op.kind != K_LS && op.kind != K_LT) return QOB_OK;
    if (op.kind == K_LS && op.terms.empty()) return QOB_OK;
    void *p;
    if (g_qubit_enable) {
        std::shared_ptr<QubitPlan> P;
        QOB_TRY(qubit_plan(op, false, dt, st, P, true));
        if (P->ok && P->hi_off < 0) {
            const int tb = std::min(P->tbits, P->nbits);
            const int nb = qubit_dot_blocks((int64_t)1 << (P->nbits - tb));
            QOB_TRY(ws_get(ctx, st, WS_RED, (size_t)nb * sizeof(double2), p));
            QubitStateDev S;
            S.nbits = P->nbits;
            S.prog_off = 0;
            S.prog_hi_off = -1;
            S.in = v;
            S.out = nullptr;
            hipError_t e = launch_qubit_dot(dt, S, u, P->tbits, (const int32_t *)P->prog.p, P->vals.p,
                                            (double2 *)p, st);
            if (e == hipSuccess) e = launch_reduce_partials(nb, (const double2 *)p, dout, st);
            if (e != hipSuccess) return fail(QOB_EHIP, "fused expect: %s", hipGetErrorString(e));
            done = true;
            return QOB_OK;
        }
    }
    GatherPlanDev G;
    if (op.kind == K_LT) {
        if (!lt_pure_sparse(op) || op.shl != op.shr) return QOB_OK;
        std::vector<OpNode *> one{&op};
        QOB_TRY(lazy_plan(op, "lt", one, {cd(1, 0)}, false, dt, G));
    } else {
        std::vector<OpNode *> lts;
        for (auto &t : op.terms) {
            if (!lt_pure_sparse(*t) || t->shl != t->shr) return QOB_OK;
            lts.push_back(t.get());
        }
        QOB_TRY(lazy_plan(op, "ls" + std::to_string(op.version), lts, op.coef, false, dt, G, true, st));
    }
    const int nb = gather_dot_blocks(n);
    QOB_TRY(ws_get(ctx, st, WS_RED, (size_t)nb * sizeof(double2), p));
    hipError_t e = launch_gather_dot(dt, G, 0, u, v, n, (double2 *)p, st);
    if (e == hipSuccess) e = launch_reduce_partials(nb, (const double2 *)p, dout, st);
    if (e != hipSuccess) return fail(QOB_EHIP, "fused expect: %s", hipGetErrorString(e));
    done = true;
    return QOB_OK;
}

// tr(op V) for an n x n V without forming op V, when op is a SparseOperator
// (one pass over its nonzeros, src/operators_sparse.jl:36-45) or a square
// all-sparse LazyTensor / LazySum (one gathered element per diagonal term).
static qob_status fused_trace(qob_ctx ctx, hipStream_t st, qob_dtype dt, OpNode &op, const void *V, int64_t n,
                              double2 *dout, bool &done) {
    done = false;
    void *p;
    if (op.kind == K_SPARSE) {
        CscDev C;
        QOB_TRY(sparse_csc(*op.sp, op.var, dt, C));
        const int nb = csc_trace_blocks(n);
        QOB_TRY(ws_get(ctx, st, WS_RED, (size_t)nb * sizeof(double2), p));
        hipError_t e = launch_csc_trace(dt, C, V, n, (double2 *)p, st);
        if (e == hipSuccess) e = launch_reduce_partials(nb, (const double2 *)p, dout, st);
        if (e != hipSuccess) return fail(QOB_EHIP, "fused trace: %s", hipGetErrorString(e));
        done = true;
        return QOB_OK;
    }
    GatherPlanDev G;
    if (op.kind == K_LT) {
        if (!lt_pure_sparse(op) || op.shl != op.shr) return QOB_OK;
        std::vector<OpNode *> one{&op};
        QOB_TRY(lazy_plan(op, "lt", one, {cd(1, 0)}, false, dt, G));
    } else if (op.kind == K_LS && !op.terms.empty()) {
        std::vector<OpNode *> lts;
        for (auto &t : op.terms) {
            if (!lt_pure_sparse(*t) || t->shl != t->shr) return QOB_OK;
            lts.push_back(t.get());
        }
        QOB_TRY(lazy_plan(op, "ls" + std::to_string(op.version), lts, op.coef, false, dt, G, true, st));
    } else {
        return QOB_OK;
    }
    const int nb = gather_dot_blocks(n);
    QOB_TRY(ws_get(ctx, st, WS_RED, (size_t)nb * sizeof(double2), p));
    hipError_t e = launch_gather_dot(dt, G, 1, nullptr, V, n, (double2 *)p, st);
    if (e == hipSuccess) e = launch_reduce_partials(nb, (const double2 *)p, dout, st);
    if (e != hipSuccess) return fail(QOB_EHIP, "fused trace: %s", hipGetErrorString(e));
    done = true;
    return QOB_OK;
}

static qob_status scalar_to_host(qob_ctx ctx, hipStream_t st, const void *d, void *host, size_t bytes) {
    (void)ctx;
    QOB_HIP(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, st));
    QOB_HIP(hipStreamSynchronize(st));
    return QOB_OK;
}

qob_status qob_expect_ket(qob_ctx ctx, qob_dtype dt, qob_op a, const void *psi, int64_t n, qob_c64 *out, qob_stream s) {
    if (!ctx || !a || !out) return fail(QOB_EARG, "null argument");
    OpNode &op = *a->node;
    if (op.m != n || op.n != n) return fail(QOB_EDIM, "expect: operator %lldx%lld on state %lld", (long long)op.m, (long long)op.n, (long long)n);
    hipStream_t st = strm(ctx, s);
    void *d;
    QOB_TRY(scal_ws(ctx, st, d));
    bool done = false;
    if (ctx->mode == QOB_MODE_FUSED) QOB_TRY(fused_bilinear(ctx, st, dt, op, psi, psi, n, (double2 *)d, done));
    if (done) return scalar_to_host(ctx, st, d, out, 16);
    // tmp = op * psi; <psi|tmp>  (src/operators.jl:119)
    void *tmp;
    QOB_TRY(ws_get(ctx, st, WS_EXP, (size_t)n * esize(dt), tmp));
    Apply ap{ctx, st, dt, 1, 1, false};
    QOB_TRY(apply_op(ap, op, psi, tmp, cd(1, 0), cd(0, 0)));
    return qob_dot(ctx, dt, n, psi, tmp, out, s);
}

qob_status qob_expect_op(qob_ctx ctx, qob_dtype dt, qob_op a, const void *rho, int64_t n, qob_c64 *out, qob_stream s) {
    if (!ctx || !a || !out) return fail(QOB_EARG, "null argument");
    OpNode &op = *a->node;
    if (op.m != n || op.n != n) return fail(QOB_EDIM, "expect: operator %lldx%lld on state %lldx%lld", (long long)op.m, (long long)op.n, (long long)n, (long long)n);
    hipStream_t st = strm(ctx, s);
    void *d;
    QOB_TRY(scal_ws(ctx, st, d));
    bool done = false;
    if (ctx->mode == QOB_MODE_FUSED) QOB_TRY(fused_trace(ctx, st, dt, op, rho, n, (double2 *)d, done));
    if (done) return scalar_to_host(ctx, st, d, out, 16);
    // tr(op * rho) (src/operators_dense.jl:223-230)
    void *tmp;
    QOB_TRY(ws_get(ctx, st, WS_EXP, (size_t)n * n * esize(dt), tmp));
    Apply ap{ctx, st, dt, 1, n, false};
    QOB_TRY(apply_op(ap, op, rho, tmp, cd(1, 0), cd(0, 0)));
    return qob_trace(ctx, dt, n, tmp, out, s);
}

qob_status qob_variance_ket(qob_ctx ctx, qob_dtype dt, qob_op a, const void *psi, int64_t n, qob_c64 *out,
                            qob_stream s) {
    if (!ctx || !a || !out) return fail(QOB_EARG, "null argument");
    OpNode &op = *a->node;
    if (op.m != n || op.n != n) return fail(QOB_EDIM, "variance: operator %lldx%lld on state %lld", (long long)op.m, (long long)op.n, (long long)n);
    hipStream_t st = strm(ctx, s);
    // x = op psi; psi' (op x) - (psi' x)^2  (src/operators.jl:139-143)
    void *x, *d;
    QOB_TRY(ws_get(ctx, st, WS_EXP, (size_t)n * esize(dt), x));
    QOB_TRY(scal_ws(ctx, st, d));
    Apply ap{ctx, st, dt, 1, 1, false};
    QOB_TRY(apply_op(ap, op, psi, x, cd(1, 0), cd(0, 0)));
    double2 *dd = (double2 *)d;
    {
        double2 *part;
        QOB_TRY(red_ws(ctx, st, n, part));
        hipError_t e = launch_dot(dt, n, psi, x, part, dd, st);
        if (e != hipSuccess) return fail(QOB_EHIP, "variance: %s", hipGetErrorString(e));
    }
    bool done = false;
    if (ctx->mode == QOB_MODE_FUSED) QOB_TRY(fused_bilinear(ctx, st, dt, op, psi, x, n, dd + 1, done));
    if (!done) {
        void *y;
        QOB_TRY(ws_get(ctx, st, WS_TMP1, (size_t)n * esize(dt), y));
        QOB_TRY(apply_op(ap, op, x, y, cd(1, 0), cd(0, 0)));
        double2 *part;
        QOB_TRY(red_ws(ctx, st, n, part));
        hipError_t e = launch_dot(dt, n, psi, y, part, dd + 1, st);
        if (e != hipSuccess) return fail(QOB_EHIP, "variance: %s", hipGetErrorString(e));
    }
    double2 h[2];
    QOB_TRY(scalar_to_host(ctx, st, d, h, sizeof(h)));
    const cd ex(h[0].x, h[0].y), ex2(h[1].x, h[1].y);
    const cd v = ex2 - ex * ex;
    out->re = v.real();
    out->im = v.imag();
    return QOB_OK;
}

qob_status qob_variance_op(qob_ctx ctx, qob_dtype dt, qob_op a, const void *rho, int64_t n, qob_c64 *out,
                           qob_stream s) {
    if (!ctx || !a || !out) return fail(QOB_EARG, "null argument");
    OpNode &op = *a->node;
    if (op.m != n || op.n != n)
        return fail(QOB_EDIM, "variance: operator %lldx%lld on state %lldx%lld", (long long)op.m, (long long)op.n,
                    (long long)n, (long long)n);
    hipStream_t st = strm(ctx, s);
    // x = op rho once; tr(op x) - tr(op rho)^2, both traces through expect
    void *x;
    QOB_TRY(ws_get(ctx, st, WS_VAR, (size_t)n * n * esize(dt), x));
    Apply ap{ctx, st, dt, 1, n, false};
    QOB_TRY(apply_op(ap, op, rho, x, cd(1, 0), cd(0, 0)));
    qob_c64 e1, e2;
    QOB_TRY(qob_expect_op(ctx, dt, a, rho, n, &e1, s));
    QOB_TRY(qob_expect_op(ctx, dt, a, x, n, &e2, s));
    const cd v = cdv(e2) - cdv(e1) * cdv(e1);
    out->re = v.real();
    out->im = v.imag();
    return QOB_OK;
}

}  // extern "C"

// ============================================================ batches
struct qob_batch_s {
    qob_ctx ctx = nullptr;
    qob_dtype dt = QOB_C64;
    int n = 0;
    int tbits = 12;
    int nitems = 0, nitems_hi = 0;
    int64_t total = 0;
    std::vector<int> nbits;
    std::vector<int32_t> prog_off, prog_hi_off;
    qob::DevBuf prog, vals, items, items_hi;
    QubitSplit xs{}, xs_hi{};  // per-XCD ranges of items / items_hi
    qob::DevBuf slot[QOB_BATCH_SLOTS];
    bool bound[QOB_BATCH_SLOTS] = {false, false, false, false};
};

struct qob_graph_s {
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    qob_ctx ctx = nullptr;
};

extern "C" {

qob_status qob_batch_create(qob_ctx ctx, int n, const qob_op *ops, qob_dtype dt, qob_batch *out) {
    if (!ctx || !out || n < 1 || !ops) return fail(QOB_EARG, "bad arguments");
    if (dt != QOB_C64 && dt != QOB_C32) return fail(QOB_EARG, "bad dtype");
    std::unique_ptr<qob_batch_s> B(new qob_batch_s());
    B->ctx = ctx;
    B->dt = dt;
    B->n = n;
    B->tbits = qubit_tbits(dt);
    if (!g_qubit_tbits) {
        // the largest tile (down to 2^10) that states holding most of the
        // elements fill: a state smaller than the tile leaves lanes idle
        for (int j = 0; j < n; ++j)
            if (!ops[j]) return fail(QOB_EARG, "null operator %d", j);
        while (B->tbits > 10) {
            int64_t big = 0, all = 0;
            for (int j = 0; j < n; ++j) {
                const int64_t d = ops[j]->node->n;
                all += d;
                if (d >= ((int64_t)1 << B->tbits)) big += d;
            }
            if (2 * big >= all) break;
            --B->tbits;
        }
    }
    std::vector<int32_t> prog;
    std::vector<cd> vals;
    for (int j = 0; j < n; ++j) {
        if (!ops[j]) return fail(QOB_EARG, "null operator %d", j);
        std::vector<TermSpec> specs;
        std::vector<std::vector<RowLists>> store;
        int nb = 0;
        QubitProgHost H;
        if (!qubit_terms(*ops[j]->node, false, specs, store, nb) || !compile_qubit(nb, specs, H, B->tbits))
            return fail(QOB_EUNSUPPORTED, "operator %d is not a qubit sum of monomial factors", j);
        // rebase the value offsets of this state's programs into the shared pool
        const int32_t vb = (int32_t)vals.size();
        rebase_qubit_prog(H.prog, vb);
        rebase_qubit_prog(H.prog_hi, vb);
        B->prog_off.push_back((int32_t)prog.size());
        prog.insert(prog.end(), H.prog.begin(), H.prog.end());
        if (!H.prog_hi.empty()) {
            B->prog_hi_off.push_back((int32_t)prog.size());
            prog.insert(prog.end(), H.prog_hi.begin(), H.prog_hi.end());
        } else {
            B->prog_hi_off.push_back(-1);
        }
        vals.insert(vals.end(), H.vals.begin(), H.vals.end());
        B->nbits.push_back(nb);
        B->total += (int64_t)1 << nb;
    }
    // Work items (state, tile) grouped by XCD.  A state's tiles read each
    // other as partners of high-bit flips, so each state goes to one XCD whole
    // (its partner tiles then sit in that XCD's L2); a state of more than
    // 1/16 of the batch's tiles is cut into power-of-two runs of tiles first.
    // Runs go largest first to the XCD with the fewest elements so far; each
    // XCD walks its runs in that order and a run's tiles in index order.
    std::vector<int> order(n);
    for (int j = 0; j < n; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return B->nbits[a] > B->nbits[b]; });
    int64_t ntiles_all = 0;
    for (int j = 0; j < n; ++j) ntiles_all += (int64_t)1 << (B->nbits[j] - std::min(B->tbits, B->nbits[j]));
    if (ntiles_all >= (1ll << 31)) return fail(QOB_EUNSUPPORTED, "batch too large");
    int64_t run_cap = 1;
    while (run_cap * 2 <= std::max<int64_t>(1, ntiles_all / 16)) run_cap *= 2;
    std::vector<QubitItemDev> xl[8], xl_hi[8];
    int64_t load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j : order) {
        const int tb = std::min(B->tbits, B->nbits[j]);
        const int64_t nt = (int64_t)1 << (B->nbits[j] - tb);
        const int64_t run = std::min(nt, run_cap);
        for (int64_t t0 = 0; t0 < nt; t0 += run) {
            const int x = (int)(std::min_element(load, load + 8) - load);
            load[x] += run << tb;
            for (int64_t k = 0; k < run; ++k) {
                // run is a power of two and t0 a multiple of it: any order of
                // the index bits below run stays inside the run
                const int64_t q = g_qubit_order == 1   ? k ^ (k >> 1)
                                  : g_qubit_order == 2 ? (((k >> 6) ^ (k >> 7)) << 6) | (k & 63)
                                                       : k;
                const int64_t t = t0 + q;
                xl[x].push_back({j, (int32_t)t});
                if (B->prog_hi_off[j] >= 0) xl_hi[x].push_back({j, (int32_t)t});
            }
        }
    }
    std::vector<QubitItemDev> items, items_hi;
    for (int x = 0; x < 8; ++x) {
        B->xs.off[x] = (int32_t)items.size();
        B->xs_hi.off[x] = (int32_t)items_hi.size();
        items.insert(items.end(), xl[x].begin(), xl[x].end());
        items_hi.insert(items_hi.end(), xl_hi[x].begin(), xl_hi[x].end());
    }
    B->xs.off[8] = (int32_t)items.size();
    B->xs_hi.off[8] = (int32_t)items_hi.size();
    B->nitems = (int)items.size();
    B->nitems_hi = (int)items_hi.size();
    if (!items_hi.empty()) QOB_TRY(B->items_hi.upload(items_hi.data(), items_hi.size() * sizeof(QubitItemDev)));
    QOB_TRY(B->prog.upload(prog.data(), prog.size() * 4));
    QOB_TRY(upload_vals(B->vals, vals, dt));
    QOB_TRY(B->items.upload(items.data(), items.size() * sizeof(QubitItemDev)));
    *out = B.release();
    return QOB_OK;
}

qob_status qob_batch_bind(qob_ctx ctx, qob_batch b, int slot, void *const *in, void *const *out) {
    if (!ctx || !b || !in || !out) return fail(QOB_EARG, "null argument");
    if (slot < 0 || slot >= QOB_BATCH_SLOTS) return fail(QOB_EARG, "slot %d out of range", slot);
    std::vector<QubitStateDev> S(b->n);
    for (int j = 0; j < b->n; ++j) {
        if (!in[j] || !out[j]) return fail(QOB_EARG, "null buffer for state %d", j);
        if (in[j] == out[j]) return fail(QOB_EALIAS, "state %d: output aliased with input", j);
        S[j].nbits = b->nbits[j];
        S[j].prog_off = b->prog_off[j];
        S[j].prog_hi_off = b->prog_hi_off[j];
        S[j].in = in[j];
        S[j].out = out[j];
    }
    // A slot bound before may still be read by an apply in flight on any stream: wait for the whole
    // device before the table is overwritten. A first bind has no readers (the copy below is
    // synchronous with the host).
    if (b->bound[slot]) QOB_HIP(hipDeviceSynchronize());
    QOB_TRY(b->slot[slot].upload(S.data(), S.size() * sizeof(QubitStateDev)));
    b->bound[slot] = true;
    return QOB_OK;
}

qob_status qob_batch_apply(qob_ctx ctx, qob_batch b, int slot, qob_c64 alpha, qob_c64 beta, qob_c64 gamma,
                           qob_stream s) {
    if (!ctx || !b) return fail(QOB_EARG, "null argument");
    if (slot < 0 || slot >= QOB_BATCH_SLOTS || !b->bound[slot]) return fail(QOB_EARG, "slot %d not bound", slot);
    hipError_t e = hipSuccess;
    if (b->nitems_hi > 0)  // high-bit flips first: out = alpha * high + beta * out
        e = launch_qubit_hi(b->dt, (const QubitStateDev *)b->slot[slot].p, (const QubitItemDev *)b->items_hi.p,
                            b->nitems_hi, b->tbits, (const int32_t *)b->prog.p, b->vals.p, d2(alpha), d2(beta),
                            strm(ctx, s), &b->xs_hi);
    if (e == hipSuccess)
        e = launch_qubit(b->dt, (const QubitStateDev *)b->slot[slot].p, (const QubitItemDev *)b->items.p, b->nitems,
                         b->tbits, (const int32_t *)b->prog.p, b->vals.p, d2(alpha), d2(beta), d2(gamma),
                         strm(ctx, s), &b->xs);
    if (e != hipSuccess) return fail(QOB_EHIP, "batch apply: %s", hipGetErrorString(e));
    return QOB_OK;
}

qob_status qob_batch_axpy(qob_ctx ctx, qob_batch b, int slot, qob_c64 a, qob_stream s) {
    if (!ctx || !b) return fail(QOB_EARG, "null argument");
    if (slot < 0 || slot >= QOB_BATCH_SLOTS || !b->bound[slot]) return fail(QOB_EARG, "slot %d not bound", slot);
    hipError_t e = launch_qubit_axpy(b->dt, (const QubitStateDev *)b->slot[slot].p, (const QubitItemDev *)b->items.p,
                                     b->nitems, b->tbits, d2(a), strm(ctx, s));
    if (e != hipSuccess) return fail(QOB_EHIP, "batch axpy: %s", hipGetErrorString(e));
    return QOB_OK;
}

qob_status qob_batch_info(qob_batch b, int64_t *total_elements, int *nstates) {
    if (!b) return fail(QOB_EARG, "null batch");
    if (total_elements) *total_elements = b->total;
    if (nstates) *nstates = b->n;
    return QOB_OK;
}

qob_status qob_batch_destroy(qob_batch b) {
    if (!b) return QOB_OK;
    if (!b->ctx) {
        delete b;
        return QOB_OK;
    }
    // applies may be queued on any stream, or captured in a live graph: the
    // batch's tables are retired until that work is done and no graph is alive
    retire(b->ctx, {std::shared_ptr<void>(b, [](void *p) { delete (qob_batch_s *)p; })});
    return QOB_OK;
}

qob_status qob_graph_begin(qob_ctx ctx, qob_stream s) {
    if (!ctx) return fail(QOB_EARG, "null context");
    QOB_HIP(hipStreamBeginCapture(strm(ctx, s), hipStreamCaptureModeThreadLocal));
    return QOB_OK;
}

qob_status qob_graph_end(qob_ctx ctx, qob_stream s, qob_graph *out) {
    if (!ctx || !out) return fail(QOB_EARG, "null argument");
    std::unique_ptr<qob_graph_s> G(new qob_graph_s());
    QOB_HIP(hipStreamEndCapture(strm(ctx, s), &G->g));
    if (hipGraphInstantiate(&G->ex, G->g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(G->g);
        return fail(QOB_EHIP, "hipGraphInstantiate failed");
    }
    G->ctx = ctx;
    {
        // from now on nothing the graph may use is freed (see retire)
        std::lock_guard<std::mutex> g(ctx->rmu);
        ctx->live_graphs++;
    }
    *out = G.release();
    return QOB_OK;
}

qob_status qob_graph_launch(qob_ctx ctx, qob_graph g, qob_stream s) {
    if (!ctx || !g) return fail(QOB_EARG, "null argument");
    QOB_HIP(hipGraphLaunch(g->ex, strm(ctx, s)));
    return QOB_OK;
}

qob_status qob_graph_destroy(qob_graph g) {
    if (!g) return QOB_OK;
    if (g->ex) (void)hipGraphExecDestroy(g->ex);
    if (g->g) (void)hipGraphDestroy(g->g);
    if (g->ctx) {
        qob_ctx ctx = g->ctx;
        std::lock_guard<std::mutex> lk(ctx->rmu);
        if (ctx->live_graphs > 0 && --ctx->live_graphs == 0 && !ctx->retired.empty()) {
            // replays may still be in flight on any stream: wait for them, then
            // free what was kept for the graphs
            const bool synced = hipDeviceSynchronize() == hipSuccess;
            reap_locked(ctx, synced);
        }
    }
    delete g;
    return QOB_OK;
}

}  // extern "C"
