/*
 * qob_hip.h — C ABI of libqob_hip.so: the operator-application hot path of
 * QuantumOpticsBase.jl (5-argument `mul!(result, op, state, α, β)`) on one
 * MI355X (gfx950).
 *
 * Every entry point is plain C: device pointers, sizes, handles and a status
 * code.  A Julia caller binds them with `ccall` (the binding lives in
 * quantumopticsbase.jl_amd/julia/QobHIP.jl); the Python mirror binds them with
 * ctypes.  Matrices are column-major, complex numbers are interleaved
 * (re, im) — the memory layout of Julia's Vector/Matrix{ComplexF64} and
 * {ComplexF32}.  Site 1 of a composite basis is the fastest-varying index
 * (reference: src/operators_dense.jl:134, src/states.jl:105 `kron(b, a)`).
 *
 * Semantics follow the reference `mul!` methods cited per entry point:
 *   result = alpha * op * state + beta * result
 * with beta == 0 overwriting result (NaNs in result are discarded) and
 * alpha == 0 for lazy operators reducing to result = beta * result
 * (src/operators_lazysum.jl:179-186, src/operators_lazytensor.jl:540).
 *
 * Thread safety: calls on one context may come from several host threads as
 * long as each uses its own stream; temporaries are kept per stream.
 */
#ifndef QOB_HIP_H
#define QOB_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ComplexF64 passed by value: same layout as Julia's ComplexF64. */
typedef struct { double re, im; } qob_c64;

/* Status codes.  The Julia binding turns each into the exception the
 * reference throws at the same place. */
typedef enum {
    QOB_OK = 0,
    QOB_EDIM = 1,          /* DimensionMismatch  (src/sparsematrix.jl:100-102,
                              src/operators_lazytensor.jl:695-703) */
    QOB_EALIAS = 2,        /* ArgumentError: output aliased with an input
                              (src/operators_lazytensor.jl:704-708) */
    QOB_EBASIS = 3,        /* IncompatibleBases (src/operators_lazysum.jl:6-11) */
    QOB_EUNSUPPORTED = 4,  /* ArgumentError / MethodError for an operand kind
                              the path does not implement
                              (src/operators_lazytensor.jl:640,677;
                               src/sparsematrix.jl:177-188) */
    QOB_ENOMEM = 5,        /* device allocation failed */
    QOB_EHIP = 6,          /* a HIP runtime call failed */
    QOB_EARG = 7           /* invalid argument (null handle, bad index, ...) */
} qob_status;

/* Element type of states and of the arithmetic.
 *   QOB_C64: ComplexF64 states, FP64 arithmetic.
 *   QOB_C32: ComplexF32 states, FP32 arithmetic (dense factors on FP32 MFMA
 *            with FP32 accumulation); reductions still accumulate in FP64. */
typedef enum { QOB_C64 = 0, QOB_C32 = 1 } qob_dtype;

/* Execution mode of a context.
 *   QOB_MODE_FUSED: LazySum terms fused into one pass per term group (default).
 *   QOB_MODE_PLAIN: the reference's loop structure — one launch per LazySum
 *                   term, each a full read-modify-write of the result
 *                   (src/operators_lazysum.jl:194-197).  Kept as the
 *                   per-call GPU baseline that the fused path is compared to. */
typedef enum { QOB_MODE_FUSED = 0, QOB_MODE_PLAIN = 1 } qob_mode;

typedef struct qob_ctx_s *qob_ctx;   /* device, default stream, workspaces */
typedef struct qob_op_s *qob_op;     /* an uploaded operator */
typedef void *qob_stream;            /* a hipStream_t; NULL = context stream */

/* ---------------------------------------------------------------- context */
qob_status qob_ctx_create(int device, qob_ctx *out);
qob_status qob_ctx_destroy(qob_ctx ctx);
qob_status qob_ctx_set_mode(qob_ctx ctx, qob_mode mode);
qob_status qob_ctx_get_stream(qob_ctx ctx, qob_stream *out);
/* Text of the last failing call on this host thread ("" if none). */
const char *qob_last_error(void);
/* Library version string. */
const char *qob_version(void);

/* ----------------------------------------------------------------- memory
 * For callers that have no device allocator of their own (the Python
 * harness); a Julia caller passes ROCArray pointers instead. */
qob_status qob_dmalloc(qob_ctx ctx, size_t bytes, void **dptr);
qob_status qob_dfree(qob_ctx ctx, void *dptr);
qob_status qob_memcpy_htod(qob_ctx ctx, void *dst, const void *src, size_t bytes, qob_stream s);
qob_status qob_memcpy_dtoh(qob_ctx ctx, void *dst, const void *src, size_t bytes, qob_stream s);
qob_status qob_memcpy_dtod(qob_ctx ctx, void *dst, const void *src, size_t bytes, qob_stream s);
qob_status qob_stream_create(qob_ctx ctx, qob_stream *out);
qob_status qob_stream_destroy(qob_ctx ctx, qob_stream s);
qob_status qob_stream_synchronize(qob_ctx ctx, qob_stream s);

/* Events on a stream: completion polling and elapsed time, for callers that
 * overlap several streams (the serving loop).  No reference counterpart:
 * the reference's mul! is synchronous on the caller's task. */
typedef void *qob_event;
qob_status qob_event_create(qob_ctx ctx, qob_event *out);
qob_status qob_event_record(qob_ctx ctx, qob_event ev, qob_stream s);
/* *done = 1 when all work recorded before the event has finished, else 0 */
qob_status qob_event_query(qob_event ev, int *done);
qob_status qob_event_synchronize(qob_event ev);
/* milliseconds between two recorded, completed events */
qob_status qob_event_elapsed(qob_event start, qob_event stop, float *ms);
qob_status qob_event_destroy(qob_event ev);

/* -------------------------------------------------------------- operators
 * Handles are reference counted: a LazyTensor or LazySum keeps its factors
 * alive; qob_op_release drops the caller's reference. */

/* SparseOperator data = SparseMatrixCSC exactly as Julia holds it:
 * 1-based Int64 colptr (n+1) and rowval (nnz), ComplexF64 nzval (nnz).
 * Replaces Operator{BL,BR,<:SparseMatrixCSC} (src/operators_sparse.jl:5,16-17). */
qob_status qob_sparse_create(qob_ctx ctx, int64_t m, int64_t n,
                             const int64_t *colptr1, const int64_t *rowval1,
                             const qob_c64 *nzval, qob_op *out);
/* DenseOperator data: column-major m x n ComplexF64 host matrix
 * (src/operators_dense.jl:49-52, 59-60). */
qob_status qob_dense_create(qob_ctx ctx, int64_t m, int64_t n,
                            const qob_c64 *data, qob_op *out);
/* identityoperator(b1, b2) = Operator(b1, b2, Eye(m, n))
 * (src/operators_sparse.jl:180-183). */
qob_status qob_eye_create(qob_ctx ctx, int64_t m, int64_t n, qob_op *out);
/* dagger(op): Operator{..,Adjoint{..}} view — a sparse op becomes
 * SparseOpAdjType (src/operators_sparse.jl:6), a dense op its adjoint. */
qob_status qob_op_adjoint(qob_ctx ctx, qob_op a, qob_op *out);
/* LazyTensor(basis_l, basis_r, indices, operators, factor)
 * (src/operators_lazytensor.jl:15-36).  shape_l/shape_r: per-site dims of the
 * composite bases; indices1: sorted 1-based site of each factor. */
qob_status qob_lazytensor_create(qob_ctx ctx, int nsites,
                                 const int64_t *shape_l, const int64_t *shape_r,
                                 int nfactors, const int64_t *indices1,
                                 const qob_op *factors, qob_c64 factor,
                                 qob_op *out);
/* LazySum(basis_l, basis_r, factors, operators)
 * (src/operators_lazysum.jl:41-51).  nterms may be 0 (zero operator). */
qob_status qob_lazysum_create(qob_ctx ctx, int64_t m, int64_t n, int nterms,
                              const qob_c64 *factors, const qob_op *operators,
                              qob_op *out);
/* LazyProduct(operators, factor) (src/operators_lazyproduct.jl:24-58):
 * factor * A1 * A2 * ... * An, applied factor by factor through temporaries
 * (mul! :103-163); operators[i-1] must have as many columns as operators[i]
 * has rows (check_multiplicable, :4-10). */
qob_status qob_lazyproduct_create(qob_ctx ctx, int nops, const qob_op *operators, qob_c64 factor, qob_op *out);
/* Replace the coefficient vector of a LazySum in place (what
 * TimeDependentSum's set_time! does, src/time_dependent_operator.jl:162-170). */
qob_status qob_lazysum_set_factors(qob_ctx ctx, qob_op sum, const qob_c64 *factors);
/* The same from device memory, stream-ordered and usable under stream
 * capture: factors[t] = dfactors[t], nterms complex doubles in device
 * memory (update_static_coefficients!, src/time_dependent_operator.jl:286-297,
 * with the coefficient evaluation left to the caller, e.g. a row of a table
 * of coefficients over a time grid).  No plan is rebuilt: the plans' values
 * are linear in the coefficients and a refresh kernel rewrites them.
 * Supported when every term is an all-sparse LazyTensor (the fused gather
 * and qubit paths); the plain mode then refuses to apply the sum, and a
 * later qob_lazysum_set_factors makes the host coefficients current again.
 * Plans are built on first use outside capture, as for every captured call. */
qob_status qob_lazysum_set_factors_d(qob_ctx ctx, qob_op sum, const void *dfactors, qob_stream s);
qob_status qob_op_shape(qob_op a, int64_t *m, int64_t *n);
qob_status qob_op_release(qob_op a);

/* ---------------------------------------------------------------- mul!
 * result = alpha * a * b + beta * result.  b and result are device arrays of
 * the given dtype; result must not alias b (QOB_EALIAS).
 *
 * mul!(::Ket, op, ::Ket, α, β):
 *   DenseOperator  src/operators_dense.jl:395
 *   SparseOperator src/operators_sparse.jl:201 -> src/sparsematrix.jl:190-214
 *   LazyTensor     src/operators_lazytensor.jl:539-557 (+ :729-739, :652-685)
 *   LazySum        src/operators_lazysum.jl:189-200 */
qob_status qob_mul_ket(qob_ctx ctx, qob_dtype dt, qob_op a,
                       const void *b, int64_t nb, void *result, int64_t nr,
                       qob_c64 alpha, qob_c64 beta, qob_stream s);
/* mul!(::Bra, ::Bra, op, α, β): result = alpha * a * op + beta * result.
 *   DenseOperator  src/operators_dense.jl:396
 *   SparseOperator src/operators_sparse.jl:202 -> src/sparsematrix.jl:216-238
 *   LazyTensor     src/operators_lazytensor.jl:559-574 (+ :711-727, :613-648)
 *   LazySum        src/operators_lazysum.jl:202-213 */
qob_status qob_mul_bra(qob_ctx ctx, qob_dtype dt, const void *a, int64_t na,
                       qob_op b, void *result, int64_t nr,
                       qob_c64 alpha, qob_c64 beta, qob_stream s);
/* mul!(::Operator, op, ::Operator, α, β): R = alpha * a * B + beta * R with B
 * (rows_b x cols) and R (rows_r x cols) contiguous column-major.
 *   DenseOperator  src/operators_dense.jl:394
 *   SparseOperator src/operators_sparse.jl:199 -> src/sparsematrix.jl:99-113,148-163
 *   LazyTensor     src/operators_lazytensor.jl:576-591 (+ :443-488, :729-739)
 *   LazySum        src/operators_lazysum.jl:215-226 */
qob_status qob_mul_op_left(qob_ctx ctx, qob_dtype dt, qob_op a,
                           const void *B, int64_t rows_b, int64_t cols,
                           void *R, int64_t rows_r,
                           qob_c64 alpha, qob_c64 beta, qob_stream s);
/* mul!(::Operator, ::Operator, op, α, β): R = alpha * B * a + beta * R with B
 * (rows x cols_b) and R (rows x cols_r).
 *   DenseOperator  src/operators_dense.jl:394
 *   SparseOperator src/operators_sparse.jl:200 -> src/sparsematrix.jl:115-146,165-175
 *   LazyTensor     src/operators_lazytensor.jl:593-609 (+ :711-727)
 *   LazySum        src/operators_lazysum.jl:227-238 */
qob_status qob_mul_op_right(qob_ctx ctx, qob_dtype dt, const void *B,
                            int64_t rows, int64_t cols_b, qob_op a,
                            void *R, int64_t cols_r,
                            qob_c64 alpha, qob_c64 beta, qob_stream s);
/* The same four products with the state and the result of different
 * element types (dt_b: state, dt_r: result), as the reference allows
 * (a ComplexF32 Ket into a ComplexF64 result, test/test_operators_lazytensor.jl:
 * 239-253; the element type is promoted, src/operators_lazytensor.jl:32-33):
 * the product runs in ComplexF64 and a ComplexF32 result is rounded once at
 * the end.  Equal element types behave exactly as the entry points above. */
qob_status qob_mul_ket_x(qob_ctx ctx, qob_dtype dt_b, qob_op a, const void *b, int64_t nb,
                         qob_dtype dt_r, void *result, int64_t nr,
                         qob_c64 alpha, qob_c64 beta, qob_stream s);
qob_status qob_mul_bra_x(qob_ctx ctx, qob_dtype dt_a, const void *a, int64_t na, qob_op b,
                         qob_dtype dt_r, void *result, int64_t nr,
                         qob_c64 alpha, qob_c64 beta, qob_stream s);
qob_status qob_mul_op_left_x(qob_ctx ctx, qob_dtype dt_b, qob_op a,
                             const void *B, int64_t rows_b, int64_t cols,
                             qob_dtype dt_r, void *R, int64_t rows_r,
                             qob_c64 alpha, qob_c64 beta, qob_stream s);
qob_status qob_mul_op_right_x(qob_ctx ctx, qob_dtype dt_b, const void *B,
                              int64_t rows, int64_t cols_b, qob_op a,
                              qob_dtype dt_r, void *R, int64_t cols_r,
                              qob_c64 alpha, qob_c64 beta, qob_stream s);
/* Dense x dense on raw device matrices: C = alpha * op(A) * B + beta * C with
 * op(A) = A (transa = 0) or transpose(A) (transa = 1, the Bra path of
 * src/operators_dense.jl:396).  mul!(::Operator, ::Operator, ::Operator)
 * when both operands are device matrices (src/operators_dense.jl:394), and
 * with n = 1 the Ket and Bra forms of a device DenseOperator
 * (src/operators_dense.jl:395-396: B is the state, C the result). */
qob_status qob_gemm(qob_ctx ctx, qob_dtype dt, int transa, int64_t m, int64_t n,
                    int64_t k, qob_c64 alpha, const void *A, const void *B,
                    qob_c64 beta, void *C, qob_stream s);

/* ---------------------------------------------------- batches and graphs
 * No counterpart in the reference (it applies one operator per mul! call):
 * many independent Kets of different lengths 2^L_j, each with its own
 * LazySum / LazyTensor of qubit terms with monomial factors (Pauli, sigma_+-,
 * number, projectors), applied in ONE persistent launch.  Buffers are bound
 * to one of QOB_BATCH_SLOTS slots once (a device pointer table), so applies
 * are pure kernel launches and can be captured into a HIP graph. */
typedef struct qob_batch_s *qob_batch;
typedef struct qob_graph_s *qob_graph;
#define QOB_BATCH_SLOTS 4
/* QOB_EUNSUPPORTED if an operator is not a qubit sum of monomial factors. */
qob_status qob_batch_create(qob_ctx ctx, int n, const qob_op *ops, qob_dtype dt, qob_batch *out);
/* Bind in[j] / out[j] (device pointers, 2^L_j elements each) to a slot.
 * Re-binding a bound slot waits for the whole device first (applies in flight
 * on any stream may read the slot's table); a first bind does not wait. */
qob_status qob_batch_bind(qob_ctx ctx, qob_batch b, int slot, void *const *in, void *const *out);
/* For every j: out_j = alpha * op_j * in_j + gamma * in_j + beta * out_j
 * (gamma fuses the Ket update of a time step; in_j must not alias out_j). */
qob_status qob_batch_apply(qob_ctx ctx, qob_batch b, int slot, qob_c64 alpha, qob_c64 beta, qob_c64 gamma,
                           qob_stream s);
/* For every j: out_j += a * in_j (the axpy of a time step). */
qob_status qob_batch_axpy(qob_ctx ctx, qob_batch b, int slot, qob_c64 a, qob_stream s);
/* Total number of basis states over the batch, and the number of states. */
qob_status qob_batch_info(qob_batch b, int64_t *total_elements, int *nstates);
qob_status qob_batch_destroy(qob_batch b);
/* Stream capture of any sequence of library calls into a HIP graph.  Every
 * captured call must have run once before (plans and workspaces exist). */
qob_status qob_graph_begin(qob_ctx ctx, qob_stream s);
qob_status qob_graph_end(qob_ctx ctx, qob_stream s, qob_graph *out);
qob_status qob_graph_launch(qob_ctx ctx, qob_graph g, qob_stream s);
qob_status qob_graph_destroy(qob_graph g);

/* ------------------------------------------------------------- tuning
 * Launch-shape knobs read by the kernels (no counterpart in the reference:
 * the reference has no device launch shapes).  Names: "sell_variant",
 * "sell_cj", "sell_jchunk", "sell_unr", "csc_rb", "csc_cc", "csc_remap",
 * "csc_rows", "csc_order", "qubit_tbits", "qubit_nth", "qubit_grid",
 * "qubit_remap", "qubit_ntfar", "qubit_order", "qubit_dbuf", "qubit_pair", "qubit_split", "qubit_enable", "mode_kernel", "fused_reg", "fused_gpw",
 * "fused_lt", "fused_tiles", "fused_resident", "nt_store", "copy_variant",
 * "sell_align", "sell_closed", "sell_cline", "sell_cpair", "sell_cjchunk", "sell_cremap", "csc_groups",
 * "csc_gkernel", "csc_gremap", "csc_waves", "csc_sweep", "ls_materialize".
 * The project's
 * autotuner (qobhip/tune.py) sweeps them and persists a table; unknown names
 * return QOB_EARG.  Settings are process-wide. */
qob_status qob_tune_set(const char *name, int64_t value);
qob_status qob_tune_get(const char *name, int64_t *value);

/* Column traversal order the right product R = B X uses for a sparse X
 * (no counterpart in the reference; host-only, needs no device).  X is given
 * as CSC (colptr[n + 1], rowind[nnz], rows 0..xrows-1).  On return *used is 1
 * and order[0..n) is the reverse Cuthill-McKee order of the pattern of X^T X
 * when the library sweeps columns in that order, else *used is 0 and order
 * is the identity (the natural order is kept when it already places columns
 * that share a row of X within 64 positions, or the reordering does not at
 * least halve that distance). */
qob_status qob_csc_column_order(int64_t n, int64_t xrows, const int32_t *colptr, const int32_t *rowind,
                                int32_t *order, int *used);

/* -------------------------------------------- elementwise and reductions
 * Reductions accumulate in FP64 for both dtypes.  The host-result variants
 * synchronise the stream; the *_d variants write one qob_c64 / double to
 * device memory and stay stream-ordered (usable under stream capture). */
/* x = a * x; a == 0 writes zeros (src/operators_lazysum.jl:179-186). */
qob_status qob_scale(qob_ctx ctx, qob_dtype dt, int64_t n, qob_c64 a, void *x, qob_stream s);
/* y = x (a plain streaming copy kernel; also the bandwidth probe). */
qob_status qob_copy(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x, void *y, qob_stream s);
/* y = a * x + y. */
qob_status qob_axpy(qob_ctx ctx, qob_dtype dt, int64_t n, qob_c64 a,
                    const void *x, void *y, qob_stream s);
/* Element-type conversion between ComplexF64 and ComplexF32 arrays. */
qob_status qob_convert(qob_ctx ctx, qob_dtype from, const void *src,
                       qob_dtype to, void *dst, int64_t n, qob_stream s);
/* dot(x, y) = sum conj(x_i) y_i  (src/states.jl:80 with dagger). */
qob_status qob_dot(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x,
                   const void *y, qob_c64 *out, qob_stream s);
/* norm(x) (src/states.jl:89). */
qob_status qob_norm(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x,
                    double *out, qob_stream s);
/* tr(A) of an n x n matrix (src/operators_dense.jl:189). */
qob_status qob_trace(qob_ctx ctx, qob_dtype dt, int64_t n, const void *A,
                     qob_c64 *out, qob_stream s);
/* normalize!(ket) (src/states.jl:90) and normalize!(op) = op / tr(op)
 * (src/operators_dense.jl:217), stream-ordered, no host round trip. */
qob_status qob_normalize_ket(qob_ctx ctx, qob_dtype dt, int64_t n, void *x, qob_stream s);
qob_status qob_normalize_op(qob_ctx ctx, qob_dtype dt, int64_t n, void *A, qob_stream s);
/* normalize!(ket) exactly as qob_normalize_ket (same kernels, bitwise the
 * same result), but allocating nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The norm's partial sums and the norm live
 * in a fixed 16 KiB + 16 B device array per GPU, held until the process
 * ends; every n fits it (the partial count is capped at 1024).  Calls that
 * overlap on one GPU share that array and race: issue them on one stream,
 * or order them with an event.  The device's array is looked up when a
 * context is created on it. */
qob_status qob_normalize_ket_noalloc(qob_ctx ctx, qob_dtype dt, int64_t n, void *x, qob_stream s);
/* expect(op, ket) = <psi|op|psi> (src/operators_dense.jl:219-221,
 * src/operators.jl:119) and expect(op, rho) = tr(op * rho)
 * (src/operators_dense.jl:223-230, src/operators_sparse.jl:36-45).  In the
 * fused mode these do not form op * state: a qubit sum or an all-sparse lazy
 * operator reduces in the apply's own pass over a Ket, a SparseOperator
 * reads only the entries of rho its nonzeros meet, an all-sparse lazy
 * operator only the entries its diagonal terms meet. */
qob_status qob_expect_ket(qob_ctx ctx, qob_dtype dt, qob_op a, const void *psi,
                          int64_t n, qob_c64 *out, qob_stream s);
qob_status qob_expect_op(qob_ctx ctx, qob_dtype dt, qob_op a, const void *rho,
                         int64_t n, qob_c64 *out, qob_stream s);
/* variance(op, ket) = <psi|op op|psi> - <psi|op|psi>^2 with x = op psi
 * computed once (src/operators.jl:139-143); FP64 result for both element
 * types.  In the fused mode <psi|op x> is a single pass when op is a qubit
 * sum or an all-sparse lazy operator. */
qob_status qob_variance_ket(qob_ctx ctx, qob_dtype dt, qob_op a, const void *psi,
                            int64_t n, qob_c64 *out, qob_stream s);
/* variance(op, rho) = tr(op op rho) - tr(op rho)^2 with op * rho formed once
 * into the library's workspace (the operator form of src/operators.jl:132-143);
 * FP64 result for both element types. */
qob_status qob_variance_op(qob_ctx ctx, qob_dtype dt, qob_op a, const void *rho,
                           int64_t n, qob_c64 *out, qob_stream s);
qob_status qob_dot_d(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x,
                     const void *y, void *dout, qob_stream s);
qob_status qob_norm_d(qob_ctx ctx, qob_dtype dt, int64_t n, const void *x,
                      void *dout, qob_stream s);
qob_status qob_trace_d(qob_ctx ctx, qob_dtype dt, int64_t n, const void *A,
                       void *dout, qob_stream s);

#ifdef __cplusplus
}
#endif

#endif /* QOB_HIP_H */
