/* Drives libqob_hip.so through include/qob_hip.h alone, from C (the way a
 * Julia ccall or any other foreign-function caller binds it):
 *   1. a SparseOperator handed over as 1-based Int64 CSC, applied left and
 *      right to a density matrix (mul!(R, S, B, a, b) and mul!(R, B, S, a, b));
 *   2. a LazySum of LazyTensors, H = sum_i J_i sx_i sx_{i+1} + sum_i h_i sz_i
 *      on 4 qubits (site 1 fastest), applied to a Ket;
 *   3. DimensionMismatch and aliasing reported as QOB_EDIM / QOB_EALIAS with
 *      a message.
 * Each result is compared with a dense product computed here.  Prints OK
 * and exits 0 when every check passes. */
#include "qob_hip.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

#define CHECK(x)                                                                                    \
    do {                                                                                            \
        qob_status st_ = (x);                                                                       \
        if (st_ != QOB_OK) {                                                                        \
            fprintf(stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #x, (int)st_, qob_last_error()); \
            return 1;                                                                               \
        }                                                                                           \
    } while (0)

static uint64_t rs = 88172645463325252ull;
static double urand(void) {
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return (double)(rs >> 11) * (1.0 / 9007199254740992.0);
}
static cplx crand(void) { return (2 * urand() - 1) + I * (2 * urand() - 1); }
static qob_c64 c64(cplx z) {
    qob_c64 r = {creal(z), cimag(z)};
    return r;
}

static double rel_err(const cplx *a, const cplx *b, size_t n) {
    double num = 0, den = 0;
    for (size_t i = 0; i < n; ++i) {
        num += pow(cabs(a[i] - b[i]), 2);
        den += pow(cabs(b[i]), 2);
    }
    return sqrt(num / (den > 0 ? den : 1));
}

static int report(const char *what, double err, double tol) {
    printf("%-34s rel err %.2e\n", what, err);
    if (!(err < tol)) {
        fprintf(stderr, "%s: error %.3e above %.1e\n", what, err, tol);
        return 1;
    }
    return 0;
}

/* one 1-based Int64 CSC from a dense column-major n x n host matrix */
static void to_csc(int n, const cplx *M, int64_t *colptr, int64_t *rowval, qob_c64 *nz) {
    int64_t p = 0;
    colptr[0] = 1;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i)
            if (M[i + (size_t)j * n] != 0) {
                rowval[p] = i + 1;
                nz[p] = c64(M[i + (size_t)j * n]);
                ++p;
            }
        colptr[j + 1] = p + 1;
    }
}

static int sparse_checks(qob_ctx ctx) {
    const int n = 50, k = 7, r = 9;
    cplx *M = calloc((size_t)n * n, sizeof(cplx));
    for (int i = 0; i < n * n; ++i)
        if (urand() < 0.1) M[i] = crand();
    int64_t *colptr = malloc((n + 1) * sizeof(int64_t)), *rowval = malloc((size_t)n * n * sizeof(int64_t));
    qob_c64 *nz = malloc((size_t)n * n * sizeof(qob_c64));
    to_csc(n, M, colptr, rowval, nz);
    qob_op S;
    CHECK(qob_sparse_create(ctx, n, n, colptr, rowval, nz, &S));
    const cplx a = 0.7 - 0.3 * I, b = 0.25 + 0.5 * I;

    /* left: R (n x k) = a M B + b R */
    cplx *B = malloc((size_t)n * k * sizeof(cplx)), *R = malloc((size_t)n * k * sizeof(cplx));
    cplx *want = malloc((size_t)n * k * sizeof(cplx)), *got = malloc((size_t)n * k * sizeof(cplx));
    for (int i = 0; i < n * k; ++i) {
        B[i] = crand();
        R[i] = crand();
    }
    for (int c = 0; c < k; ++c)
        for (int i = 0; i < n; ++i) {
            cplx s = 0;
            for (int j = 0; j < n; ++j) s += M[i + (size_t)j * n] * B[j + (size_t)c * n];
            want[i + (size_t)c * n] = a * s + b * R[i + (size_t)c * n];
        }
    void *dB, *dR;
    CHECK(qob_dmalloc(ctx, (size_t)n * k * 16, &dB));
    CHECK(qob_dmalloc(ctx, (size_t)n * k * 16, &dR));
    CHECK(qob_memcpy_htod(ctx, dB, B, (size_t)n * k * 16, NULL));
    CHECK(qob_memcpy_htod(ctx, dR, R, (size_t)n * k * 16, NULL));
    CHECK(qob_mul_op_left(ctx, QOB_C64, S, dB, n, k, dR, n, c64(a), c64(b), NULL));
    CHECK(qob_memcpy_dtoh(ctx, got, dR, (size_t)n * k * 16, NULL));
    if (report("SparseOperator * rho (left)", rel_err(got, want, (size_t)n * k), 1e-12)) return 1;

    /* right: R2 (r x n) = a B2 M + b R2 */
    cplx *B2 = malloc((size_t)r * n * sizeof(cplx)), *R2 = malloc((size_t)r * n * sizeof(cplx));
    cplx *want2 = malloc((size_t)r * n * sizeof(cplx)), *got2 = malloc((size_t)r * n * sizeof(cplx));
    for (int i = 0; i < r * n; ++i) {
        B2[i] = crand();
        R2[i] = crand();
    }
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < r; ++i) {
            cplx s = 0;
            for (int j = 0; j < n; ++j) s += B2[i + (size_t)j * r] * M[j + (size_t)c * n];
            want2[i + (size_t)c * r] = a * s + b * R2[i + (size_t)c * r];
        }
    void *dB2, *dR2;
    CHECK(qob_dmalloc(ctx, (size_t)r * n * 16, &dB2));
    CHECK(qob_dmalloc(ctx, (size_t)r * n * 16, &dR2));
    CHECK(qob_memcpy_htod(ctx, dB2, B2, (size_t)r * n * 16, NULL));
    CHECK(qob_memcpy_htod(ctx, dR2, R2, (size_t)r * n * 16, NULL));
    CHECK(qob_mul_op_right(ctx, QOB_C64, dB2, r, n, S, dR2, n, c64(a), c64(b), NULL));
    CHECK(qob_memcpy_dtoh(ctx, got2, dR2, (size_t)r * n * 16, NULL));
    if (report("rho * SparseOperator (right)", rel_err(got2, want2, (size_t)r * n), 1e-12)) return 1;

    /* errors: wrong inner dimension, aliased output */
    qob_status st = qob_mul_op_left(ctx, QOB_C64, S, dB, n - 1, k, dR, n, c64(a), c64(b), NULL);
    if (st != QOB_EDIM || strlen(qob_last_error()) == 0) {
        fprintf(stderr, "dimension mismatch gave %d\n", (int)st);
        return 1;
    }
    st = qob_mul_op_left(ctx, QOB_C64, S, dB, n, k, dB, n, c64(a), c64(b), NULL);
    if (st != QOB_EALIAS || strlen(qob_last_error()) == 0) {
        fprintf(stderr, "aliasing gave %d\n", (int)st);
        return 1;
    }
    printf("%-34s %s\n", "DimensionMismatch / aliasing", "reported");
    CHECK(qob_dfree(ctx, dB));
    CHECK(qob_dfree(ctx, dR));
    CHECK(qob_dfree(ctx, dB2));
    CHECK(qob_dfree(ctx, dR2));
    CHECK(qob_op_release(S));
    free(M), free(colptr), free(rowval), free(nz), free(B), free(R), free(want), free(got);
    free(B2), free(R2), free(want2), free(got2);
    return 0;
}

static int lazysum_check(qob_ctx ctx) {
    enum { L = 4, D = 16 };
    /* sigma_x, sigma_z as 2 x 2 CSC, 1-based */
    const int64_t cp_x[3] = {1, 2, 3}, rv_x[2] = {2, 1};
    const qob_c64 nz_x[2] = {{1, 0}, {1, 0}};
    const int64_t cp_z[3] = {1, 2, 3}, rv_z[2] = {1, 2};
    const qob_c64 nz_z[2] = {{1, 0}, {-1, 0}};
    qob_op sx, sz;
    CHECK(qob_sparse_create(ctx, 2, 2, cp_x, rv_x, nz_x, &sx));
    CHECK(qob_sparse_create(ctx, 2, 2, cp_z, rv_z, nz_z, &sz));
    const int64_t shape[L] = {2, 2, 2, 2};
    qob_op terms[2 * L - 1];
    qob_c64 coef[2 * L - 1];
    double J[L - 1], h[L];
    const qob_c64 one = {1, 0};
    for (int i = 0; i < L - 1; ++i) {
        J[i] = 2 * urand() - 1;
        const int64_t idx[2] = {i + 1, i + 2};
        const qob_op f[2] = {sx, sx};
        CHECK(qob_lazytensor_create(ctx, L, shape, shape, 2, idx, f, one, &terms[i]));
        coef[i] = c64(J[i]);
    }
    for (int i = 0; i < L; ++i) {
        h[i] = 2 * urand() - 1;
        const int64_t idx[1] = {i + 1};
        CHECK(qob_lazytensor_create(ctx, L, shape, shape, 1, idx, &sz, one, &terms[L - 1 + i]));
        coef[L - 1 + i] = c64(h[i]);
    }
    qob_op H;
    CHECK(qob_lazysum_create(ctx, D, D, 2 * L - 1, coef, terms, &H));
    /* dense H: basis state s, bit i of s = site i + 1 (site 1 fastest); sz = +1 on bit 0 */
    cplx Hd[D * D];
    memset(Hd, 0, sizeof(Hd));
    for (int s = 0; s < D; ++s) {
        for (int i = 0; i < L; ++i) Hd[s + s * D] += h[i] * (((s >> i) & 1) ? -1.0 : 1.0);
        for (int i = 0; i < L - 1; ++i) {
            const int t = s ^ (1 << i) ^ (1 << (i + 1));
            Hd[t + s * D] += J[i];
        }
    }
    cplx psi[D], out0[D], want[D], got[D];
    for (int s = 0; s < D; ++s) {
        psi[s] = crand();
        out0[s] = NAN;  /* beta = 0 overwrites */
    }
    for (int t = 0; t < D; ++t) {
        cplx acc = 0;
        for (int s = 0; s < D; ++s) acc += Hd[t + s * D] * psi[s];
        want[t] = 1.5 * acc;
    }
    void *dpsi, *dout;
    CHECK(qob_dmalloc(ctx, D * 16, &dpsi));
    CHECK(qob_dmalloc(ctx, D * 16, &dout));
    CHECK(qob_memcpy_htod(ctx, dpsi, psi, D * 16, NULL));
    CHECK(qob_memcpy_htod(ctx, dout, out0, D * 16, NULL));
    CHECK(qob_mul_ket(ctx, QOB_C64, H, dpsi, D, dout, D, c64(1.5), c64(0), NULL));
    CHECK(qob_memcpy_dtoh(ctx, got, dout, D * 16, NULL));
    if (report("LazySum of LazyTensors * Ket", rel_err(got, want, D), 1e-12)) return 1;
    CHECK(qob_dfree(ctx, dpsi));
    CHECK(qob_dfree(ctx, dout));
    for (int i = 0; i < 2 * L - 1; ++i) CHECK(qob_op_release(terms[i]));
    CHECK(qob_op_release(H));
    CHECK(qob_op_release(sx));
    CHECK(qob_op_release(sz));
    return 0;
}

int main(void) {
    printf("%s\n", qob_version());
    qob_ctx ctx;
    CHECK(qob_ctx_create(0, &ctx));
    if (sparse_checks(ctx) || lazysum_check(ctx)) return 1;
    CHECK(qob_ctx_destroy(ctx));
    printf("OK\n");
    return 0;
}
