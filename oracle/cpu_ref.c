/*
 * ORACLE / CPU BASELINE — C restatement of the reference's sparse x dense
 * loops, for timing on the host cores and as a full-size checker.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and by bench.py's cpu_baseline
 * leg; never by the product library.
 *
 *   qobref_gemm_sp_dense  gemm!(alpha, M::SparseMatrixCSC, B, beta, result)
 *                         src/sparsematrix.jl:99-113 with the nnz > 550 loop
 *                         gemm_sp_dense_big :25-47 (j outer, CSC inner)
 *   qobref_gemm_dense_sp  gemm!(alpha, B, M::SparseMatrixCSC, beta, result)
 *                         src/sparsematrix.jl:115-146 (column axpys)
 *   qobref_lazysum_ket    mul!(result::Ket, H::LazySum, psi::Ket, alpha, beta)
 *                         for LazySums of all-sparse square LazyTensors:
 *                         src/operators_lazysum.jl:189-200 (term 1 takes beta,
 *                         the others beta = 1), each term through
 *                         _gemm_puresparse (src/operators_lazytensor.jl:729-739)
 *                         and the recursion _gemm_recursive_lazy_dense (:652-685)
 *   qobref_lazysum_ket_flat  the same product with each term's recursion
 *                         unrolled into nested loops (equal bit for bit)
 * Complex numbers are interleaved doubles, matrices column-major, CSC index
 * arrays 1-based int64 exactly as Julia holds them.  Single-threaded, like the
 * reference's hand-written loops, except the _omp variants of the two sparse
 * products (an all-cores upper bound).
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

static void beta_prep(cplx beta, cplx *result, int64_t n) {
    if (beta == 0) {
        memset(result, 0, (size_t)n * sizeof(cplx));
    } else if (beta != 1) {
        for (int64_t i = 0; i < n; ++i) result[i] *= beta;
    }
}

void qobref_gemm_sp_dense(double alpha_re, double alpha_im, int64_t m, int64_t n, const int64_t *colptr,
                          const int64_t *rowval, const cplx *nzval, const cplx *B, int64_t ncols,
                          double beta_re, double beta_im, cplx *result) {
    const cplx alpha = alpha_re + alpha_im * I, beta = beta_re + beta_im * I;
    beta_prep(beta, result, m * ncols);
    const int one = (alpha == 1);
    for (int64_t j = 0; j < ncols; ++j) {
        const cplx *bj = B + j * n;
        cplx *rj = result + j * m;
        for (int64_t c = 0; c < n; ++c) {
            const cplx m2 = one ? bj[c] : alpha * bj[c];
            for (int64_t i = colptr[c] - 1; i < colptr[c + 1] - 1; ++i) rj[rowval[i] - 1] += nzval[i] * m2;
        }
    }
}

void qobref_gemm_dense_sp(double alpha_re, double alpha_im, int64_t rows, int64_t m, int64_t n,
                          const int64_t *colptr, const int64_t *rowval, const cplx *nzval, const cplx *B,
                          double beta_re, double beta_im, cplx *result) {
    const cplx alpha = alpha_re + alpha_im * I, beta = beta_re + beta_im * I;
    (void)m;
    beta_prep(beta, result, rows * n);
    const int one = (alpha == 1);
    for (int64_t c = 0; c < n; ++c) {
        cplx *rc = result + c * rows;
        for (int64_t i = colptr[c] - 1; i < colptr[c + 1] - 1; ++i) {
            const cplx mi = one ? nzval[i] : nzval[i] * alpha;
            const cplx *bcol = B + (rowval[i] - 1) * rows;
            for (int64_t j = 0; j < rows; ++j) rc[j] += mi * bcol[j];
        }
    }
}

/* The same two loops with their outer loop over output columns (independent
 * in both: column j of the result depends on column j of B alone on the left,
 * on the columns of B that column c of M names on the right) spread over
 * nthreads OpenMP threads, beta applied per column.  An all-cores upper bound
 * for the CPU path; the reference's loops themselves are single-threaded. */
void qobref_gemm_sp_dense_omp(double alpha_re, double alpha_im, int64_t m, int64_t n, const int64_t *colptr,
                              const int64_t *rowval, const cplx *nzval, const cplx *B, int64_t ncols,
                              double beta_re, double beta_im, cplx *result, int nthreads) {
    const cplx alpha = alpha_re + alpha_im * I, beta = beta_re + beta_im * I;
    const int one = (alpha == 1);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t j = 0; j < ncols; ++j) {
        const cplx *bj = B + j * n;
        cplx *rj = result + j * m;
        beta_prep(beta, rj, m);
        for (int64_t c = 0; c < n; ++c) {
            const cplx m2 = one ? bj[c] : alpha * bj[c];
            for (int64_t i = colptr[c] - 1; i < colptr[c + 1] - 1; ++i) rj[rowval[i] - 1] += nzval[i] * m2;
        }
    }
}

void qobref_gemm_dense_sp_omp(double alpha_re, double alpha_im, int64_t rows, int64_t m, int64_t n,
                              const int64_t *colptr, const int64_t *rowval, const cplx *nzval, const cplx *B,
                              double beta_re, double beta_im, cplx *result, int nthreads) {
    const cplx alpha = alpha_re + alpha_im * I, beta = beta_re + beta_im * I;
    (void)m;
    const int one = (alpha == 1);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t c = 0; c < n; ++c) {
        cplx *rc = result + c * rows;
        beta_prep(beta, rc, rows);
        for (int64_t i = colptr[c] - 1; i < colptr[c + 1] - 1; ++i) {
            const cplx mi = one ? nzval[i] : nzval[i] * alpha;
            const cplx *bcol = B + (rowval[i] - 1) * rows;
            for (int64_t j = 0; j < rows; ++j) rc[j] += mi * bcol[j];
        }
    }
}

/* One term of a LazySum: factor f of the term sits at site fsite[f] (1-based)
 * with CSC arrays at the given offsets of the pooled arrays. */
typedef struct {
    int nsites;
    const int64_t *shape;    /* per-site dimension */
    const int64_t *strides;  /* column-major strides, site 1 fastest */
    const int *fac_of_site;  /* site (0-based) -> factor index in the pools, or -1 */
    const int64_t *fdim, *fcp_off, *fnz_off;
    const int64_t *colptr, *rowval;
    const cplx *nzval;
    const cplx *psi;
    cplx *result;
} term_ctx;

/* src/operators_lazytensor.jl:652-685 with 1-based K, J */
static void rec_lazy_dense(const term_ctx *t, int i_k, int64_t K, int64_t J, cplx val) {
    if (i_k > t->nsites) {
        t->result[J - 1] += val * t->psi[K - 1];
        return;
    }
    const int64_t st = t->strides[i_k - 1];
    const int f = t->fac_of_site[i_k - 1];
    if (f >= 0) {
        const int64_t *cp = t->colptr + t->fcp_off[f];
        const int64_t *rv = t->rowval + t->fnz_off[f];
        const cplx *nz = t->nzval + t->fnz_off[f];
        for (int64_t k = 1; k <= t->fdim[f]; ++k) {
            const int64_t K_ = K + st * (k - 1);
            for (int64_t jp = cp[k - 1]; jp < cp[k]; ++jp) {
                const int64_t j = rv[jp - 1];
                rec_lazy_dense(t, i_k + 1, K_, J + st * (j - 1), val * nz[jp - 1]);
            }
        }
        return;
    }
    for (int64_t k = 1; k <= t->shape[i_k - 1]; ++k)
        rec_lazy_dense(t, i_k + 1, K + st * (k - 1), J + st * (k - 1), val);
}

/* coef[t] = LazySum coefficient times the LazyTensor factor of term t;
 * term t owns factors term_off[t] .. term_off[t+1]-1 of the pools. */
void qobref_lazysum_ket(double alpha_re, double alpha_im, int nsites, const int64_t *shape, int nterms,
                        const double *coef, const int64_t *term_off, const int64_t *fsite, const int64_t *fdim,
                        const int64_t *fcp_off, const int64_t *fnz_off, const int64_t *colptr,
                        const int64_t *rowval, const cplx *nzval, const cplx *psi, double beta_re, double beta_im,
                        cplx *result) {
    const cplx alpha = alpha_re + alpha_im * I, beta = beta_re + beta_im * I;
    int64_t strides[64];
    int fac_of_site[64];
    int64_t D = 1;
    for (int i = 0; i < nsites && i < 64; ++i) {
        strides[i] = D;
        D *= shape[i];
    }
    if (nterms == 0) {
        beta_prep(beta, result, D);  /* empty LazySum: zero operator */
        return;
    }
    for (int t = 0; t < nterms; ++t) {
        beta_prep(t == 0 ? beta : 1, result, D);
        for (int i = 0; i < nsites; ++i) fac_of_site[i] = -1;
        for (int64_t f = term_off[t]; f < term_off[t + 1]; ++f) fac_of_site[fsite[f] - 1] = (int)f;
        term_ctx c = {nsites, shape, strides, fac_of_site, fdim, fcp_off, fnz_off, colptr, rowval, nzval, psi,
                      result};
        const cplx ct = coef[2 * t] + coef[2 * t + 1] * I;
        rec_lazy_dense(&c, 1, 1, 1, alpha * ct);
    }
}

/* qobref_lazysum_ket_flat: the same sum with each term's recursion unrolled.
 * For a fixed term the recursion of _gemm_recursive_lazy_dense
 * (src/operators_lazytensor.jl:652-685) visits every choice of
 * (column k, stored entry jp) at the factor sites, in site order, and for
 * each choice every multi-index of the sites without a factor; the leaf adds
 * val * psi[K] to result[J] with val the running product alpha * coef * nz_1
 * * nz_2 ... in site order.  Here the factor choices are listed once, in the
 * same nested order, as (J offset, K offset, val); the sites without a
 * factor, which only carry the index through, are swept as runs of
 * consecutive sites (a run starting at site s covers offsets strides[s] * m,
 * m < product of its dims), and at each point of that sweep the choices are
 * applied in their listed order.  A given result[J] therefore receives the
 * same products in the same order as in the recursion, and each product is
 * the C99 complex product written out ((a c - b d) + i (a d + b c), no FMA
 * contraction under -std=c11): for finite values the results are equal bit
 * for bit.  Test infrastructure for checking long sequences of applies at
 * full size. */
typedef struct {
    int nblk;
    int64_t bst[65], bn[65];  /* runs of sites without a factor: stride, length */
    int64_t nch;              /* factor choices */
    const int64_t *offJ, *offK;
    const double *vre, *vim;
    const double *psi;        /* interleaved re, im */
    double *result;
} flat_ctx;

static void flat_sweep(const flat_ctx *c, int b, int64_t base) {
    if (b < 0) {
        for (int64_t h = 0; h < c->nch; ++h) {
            const double *p = c->psi + 2 * (base + c->offK[h]);
            double *r = c->result + 2 * (base + c->offJ[h]);
            const double pr = p[0], pi = p[1], vr = c->vre[h], vi = c->vim[h];
            r[0] += vr * pr - vi * pi;
            r[1] += vr * pi + vi * pr;
        }
        return;
    }
    const int64_t st = c->bst[b], n = c->bn[b];
    if (b > 0) {
        for (int64_t m = 0; m < n; ++m) flat_sweep(c, b - 1, base + st * m);
        return;
    }
    const int64_t nch = c->nch;
    for (int64_t m = 0; m < n; ++m) {
        const int64_t bm = base + st * m;
        for (int64_t h = 0; h < nch; ++h) {
            const double *p = c->psi + 2 * (bm + c->offK[h]);
            double *r = c->result + 2 * (bm + c->offJ[h]);
            const double pr = p[0], pi = p[1], vr = c->vre[h], vi = c->vim[h];
            r[0] += vr * pr - vi * pi;
            r[1] += vr * pi + vi * pr;
        }
    }
}

void qobref_lazysum_ket_flat(double alpha_re, double alpha_im, int nsites, const int64_t *shape, int nterms,
                             const double *coef, const int64_t *term_off, const int64_t *fsite,
                             const int64_t *fdim, const int64_t *fcp_off, const int64_t *fnz_off,
                             const int64_t *colptr, const int64_t *rowval, const cplx *nzval, const cplx *psi,
                             double beta_re, double beta_im, cplx *result) {
    const cplx alpha = alpha_re + alpha_im * I, beta = beta_re + beta_im * I;
    int64_t strides[64];
    int64_t D = 1;
    for (int i = 0; i < nsites && i < 64; ++i) {
        strides[i] = D;
        D *= shape[i];
    }
    if (nterms == 0) {
        beta_prep(beta, result, D);
        return;
    }
    for (int t = 0; t < nterms; ++t) {
        beta_prep(t == 0 ? beta : 1, result, D);
        int fac_of_site[64];
        for (int i = 0; i < nsites; ++i) fac_of_site[i] = -1;
        for (int64_t f = term_off[t]; f < term_off[t + 1]; ++f) fac_of_site[fsite[f] - 1] = (int)f;
        flat_ctx c;
        c.nblk = 0;
        int nfac = 0;
        int64_t fst[64], fidx[64], nch = 1;
        for (int i = 0; i < nsites; ++i) {
            if (fac_of_site[i] >= 0) {
                fst[nfac] = strides[i];
                fidx[nfac++] = fac_of_site[i];
            } else if (i > 0 && fac_of_site[i - 1] < 0) {
                c.bn[c.nblk - 1] *= shape[i];
            } else {
                c.bst[c.nblk] = strides[i];
                c.bn[c.nblk++] = shape[i];
            }
        }
        for (int f = 0; f < nfac; ++f) {
            const int64_t *cp = colptr + fcp_off[fidx[f]];
            nch *= cp[fdim[fidx[f]]] - cp[0];
        }
        if (nch == 0) continue;
        int64_t *oJ = malloc(sizeof(int64_t) * 2 * nch);
        double *v = malloc(sizeof(double) * 2 * nch);
        if (!oJ || !v) {
            free(oJ);
            free(v);
            return;
        }
        /* the choices in the recursion's nested order: factor 1 outermost */
        const cplx ct = coef[2 * t] + coef[2 * t + 1] * I;
        int64_t pos[64], kk[64], J0 = 0, K0 = 0, h = 0;
        cplx val[65];
        val[0] = alpha * ct;
        int f = 0;
        if (nfac == 0) {
            oJ[0] = 0;
            oJ[nch] = 0;
            v[0] = creal(val[0]);
            v[nch] = cimag(val[0]);
            h = 1;
        } else {
            /* iterative depth-first walk over (k, jp) per factor */
            kk[0] = 1;
            pos[0] = colptr[fcp_off[fidx[0]]];
            int64_t Js[65], Ks[65];
            Js[0] = J0;
            Ks[0] = K0;
            while (f >= 0) {
                const int64_t fi = fidx[f];
                const int64_t *cp = colptr + fcp_off[fi];
                /* advance to the next stored entry of this factor, column by column */
                while (kk[f] <= fdim[fi] && pos[f] >= cp[kk[f]]) {
                    ++kk[f];
                    if (kk[f] <= fdim[fi]) pos[f] = cp[kk[f] - 1];
                }
                if (kk[f] > fdim[fi]) {
                    --f;
                    if (f >= 0) ++pos[f];
                    continue;
                }
                const int64_t jp = pos[f];
                const int64_t j = rowval[fnz_off[fi] + jp - 1];
                Js[f + 1] = Js[f] + fst[f] * (j - 1);
                Ks[f + 1] = Ks[f] + fst[f] * (kk[f] - 1);
                val[f + 1] = val[f] * nzval[fnz_off[fi] + jp - 1];
                if (f + 1 == nfac) {
                    oJ[h] = Js[f + 1];
                    oJ[nch + h] = Ks[f + 1];
                    v[h] = creal(val[f + 1]);
                    v[nch + h] = cimag(val[f + 1]);
                    ++h;
                    ++pos[f];
                } else {
                    ++f;
                    kk[f] = 1;
                    pos[f] = colptr[fcp_off[fidx[f]]];
                }
            }
        }
        c.nch = h;
        c.offJ = oJ;
        c.offK = oJ + nch;
        c.vre = v;
        c.vim = v + nch;
        c.psi = (const double *)psi;
        c.result = (double *)result;
        flat_sweep(&c, c.nblk - 1, 0);
        free(oJ);
        free(v);
    }
}
