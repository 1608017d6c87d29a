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
 * Complex numbers are interleaved doubles, matrices column-major, CSC index
 * arrays 1-based int64 exactly as Julia holds them.  Single-threaded, like the
 * reference's hand-written loops.
 */
#include <complex.h>
#include <stdint.h>
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
