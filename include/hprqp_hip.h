/*
 * Extensions of the HPR-QP C API that only the MI355X backend has. None of
 * these replace a reference symbol; they exist for callers that need to step
 * the iteration (parity tests, drift tools like the reference's
 * examples/gpu/compare_gpu_unchecked_path.c) and for the batched solver.
 */
#ifndef HPRQP_HIP_H
#define HPRQP_HIP_H

#include <stddef.h>

#include "hprqp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A device-resident solve: upload + device transpose + device scaling +
 * setup happen in create; advance runs bodies of the solve loop of
 * src/shared/solver_backend.c:432-513 (one body = residual check if due,
 * termination test, restart / sigma update, one HPR iteration, M-norm if due).
 * Vectors read back are in the scaled space unless noted. With
 * params->polishing set (the default, as in the reference) check points also
 * feed the host polishing worker when one is installed, so the model and the
 * strings params points to must outlive the session's last advance. */
typedef struct HPRQPHipSession HPRQPHipSession;

HPRQPStatus hprqp_hip_session_create(const HPRQPModel *model, const HPRQPParams *params,
                                     HPRQPHipSession **out_session);
void hprqp_hip_session_destroy(HPRQPHipSession *session);
/* Runs up to `bodies` loop bodies; *finished = 1 once the loop terminated. */
HPRQPStatus hprqp_hip_session_advance(HPRQPHipSession *session, size_t bodies, int *finished);
/* Copies the named device vector (x, x_bar, w, w_bar, y, y_bar, z_bar, s, ATy,
 * Qw, x_hat, last_x, last_w, last_y, last_ATy, dx, dy, dw, c, l, u, AL, AU,
 * row_norm, col_norm; with the equality-row metric active also gram_r, the
 * last step's E t - b of the metric rows, and gram_step, its G^{-1} solve)
 * to host; returns its length (0: unknown name). */
size_t hprqp_hip_session_read(const HPRQPHipSession *session, const char *name, double *out, size_t capacity);
/* Named scalars (sigma, lambda_max_A, lambda_max_Q, power_steps_A and
 * power_steps_Q (products taken by the two power iterations), Q_is_diag, noC, iter,
 * inner, restart_times, kkt, err_rp, err_rd, gap, pobj, dobj, norm_b_org,
 * norm_c_org, gram_active (0 off, 1 plain-QP metric, 2 row-cone prefix),
 * gram_rows, gram_beta, gram_mode); NaN if unknown. */
double hprqp_hip_session_scalar(const HPRQPHipSession *session, const char *name);
/* Scaled device CSR of "Q", "A" or "AT" copied to host (int32 indices
 * widened to size_t); returns nnz. */
size_t hprqp_hip_session_matrix(const HPRQPHipSession *session, const char *name, size_t *row_ptr,
                                size_t *col_ind, double *values);
/* Final results exactly as solve_gpu reports them (original units). */
HPRQPStatus hprqp_hip_session_results(HPRQPHipSession *session, HPRQPResults *results);

/* Times `iterations` HPR iterations (non-check variant, advancing the
 * session's iterate) with HIP events on the session's stream. For each
 * iteration kernel kind k < return value: names + 32*k (NUL-terminated),
 * mean_ms[k] per launch, bytes_per_launch[k] (algorithmic HBM bytes),
 * launches[k]. capacity >= 9 (the kinds: zxw1_q, tempv_q, y_a, w2_at,
 * zxw1_noc_at, qw_noc_q, cone_x, cone_y, gram_solve). Returns -1 on error. */
int hprqp_hip_session_profile(HPRQPHipSession *session, size_t iterations, int capacity, char *names,
                              double *mean_ms, double *bytes_per_launch, size_t *launches);

/* Kernel-level entry points on caller-provided device buffers (int32 CSR,
 * FP64 values, device pointers). stream may be NULL for the default stream.
 * y = M x with the row-blocked kernel; the row blocks are planned from the
 * row pointer on each call (device memory allocated and freed by the call),
 * which waits for the product before it returns. */
HPRQPStatus hprqp_hip_csr_spmv(int nrow, const int *row_ptr, const int *col_ind, const double *values,
                               const double *x, double *y, void *stream);
/* hprqp_hip_csr_spmv without any allocation: no hipMalloc / hipFree and no
 * heap allocation; the row-block plan lives in a fixed device array of the
 * library (4 MiB on each GPU it runs on, held until the process ends).
 * nrow <= 524288 (2^19), else INVALID_ARGUMENT with a message on stderr.
 * Same y as hprqp_hip_csr_spmv, bit for bit. Calls are serialised inside the
 * process and each waits for its product; calls from different processes
 * use different copies. No counterpart in the reference: for callers that
 * cannot allocate (tight serving loops, harnesses that forbid hipMalloc). */
HPRQPStatus hprqp_hip_csr_spmv_noalloc(int nrow, const int *row_ptr, const int *col_ind, const double *values,
                                       const double *x, double *y, void *stream);
/* Device transpose with the stable (column, row) ordering of the host
 * counting-sort transpose; perm (optional) receives, for each output entry,
 * its source entry index. */
HPRQPStatus hprqp_hip_csr_transpose(int nrow, int ncol, int nnz, const int *row_ptr, const int *col_ind,
                                    const double *values, int *t_row_ptr, int *t_col_ind, double *t_values,
                                    int *perm, void *stream);

/* Batched solver: `count` independent QPs (built with build_from_QAbc; no
 * cones) solved by one persistent launch with one
 * workgroup per instance and the whole HPR state machine on the device.
 * create uploads, scales and runs setup + the iteration-0 check on the
 * device; solve relaunches the chunk kernel (bodies_per_launch loop bodies
 * per instance per launch) until every instance terminates (stoptol,
 * max_iter from params) or params->time_limit passes; result returns
 * instance k in original units. No counterpart in the reference, which
 * solves one QP per call. */
typedef struct HPRQPHipBatch HPRQPHipBatch;
HPRQPStatus hprqp_hip_batch_create(const HPRQPModel *models, size_t count, const HPRQPParams *params,
                                   HPRQPHipBatch **out_batch);
HPRQPStatus hprqp_hip_batch_solve(HPRQPHipBatch *batch, size_t bodies_per_launch, size_t *remaining);
double hprqp_hip_batch_seconds(const HPRQPHipBatch *batch);
/* Row length from which the batch's CSR engine reduces a row with a whole
 * wavefront (below it: LDS-staged sequential row sums). Chosen at creation by
 * timing a probe of one iteration's products over 16, 32, 64, 128 and
 * "never", or fixed by the environment variable HPRQP_BATCH_WAVE_ROW; the
 * choice changes only the summation order of long rows. No counterpart in
 * the reference (its SpMV-mode autotune, src/gpu/iteration.cu:1287-1382,
 * picks between whole SpMV implementations). */
int hprqp_hip_batch_wave_row(const HPRQPHipBatch *batch);
HPRQPStatus hprqp_hip_batch_result(const HPRQPHipBatch *batch, size_t index, HPRQPResults *results);
void hprqp_hip_batch_destroy(HPRQPHipBatch *batch);
/* Batch of `count` QPs of one shape from dense row-major host arrays: Q
 * (count x n x n), A (count x m x n), c, l, u (count x n), AL, AU
 * (count x m); n, m <= 2048. Same result as build_from_QAbc on every
 * instance followed by hprqp_hip_batch_create (Q is symmetrised on the
 * device with the same drop rule), without host models; INVALID_ARGUMENT if
 * a row of A has no nonzero entry or no finite bound (build_from_QAbc would
 * delete it). No counterpart in the reference: the entry for
 * differentiable-layer callers that hold dense batched tensors. */
HPRQPStatus hprqp_hip_batch_create_dense(size_t count, size_t n, size_t m, const double *Q, const double *c,
                                         const double *A, const double *AL, const double *AU, const double *l,
                                         const double *u, const HPRQPParams *params, HPRQPHipBatch **out_batch);
/* All instances' solutions of a solved batch in original units, instances
 * concatenated: x and z (sum of n), y (sum of m); per instance status,
 * iterations, primal objective and KKT residual. Null outputs are skipped.
 * No counterpart in the reference (bulk form of hprqp_hip_batch_result). */
HPRQPStatus hprqp_hip_batch_solution(const HPRQPHipBatch *batch, double *x, double *y, double *z, int *status,
                                     size_t *iterations, double *objective, double *kkt);
/* Adjoint of a solved batch (implicit differentiation of each converged KKT
 * point, active set from the projections, the active-set KKT system solved
 * matrix-free by MINRES with the batch's SpMV kernels). dl_dx: upstream
 * gradient in original units, instances concatenated (sum of n). Outputs u
 * (sum n), v_rows (sum m), v_bounds (sum n), steps (count); host arrays.
 * With x, y, z the solution: dL/dc = -u, dL/dQ = -(u x' + x u')/2,
 * dL/dA = y u' - v_rows x', dL/d(active row bound) = v_rows,
 * dL/d(active variable bound) = v_bounds. No counterpart in the reference. */
HPRQPStatus hprqp_hip_batch_adjoint(HPRQPHipBatch *batch, const double *dl_dx, double tol, int max_steps, double *u,
                                    double *v_rows, double *v_bounds, int *steps);

/* Zero-copy forms of the three calls above for callers whose data already
 * lives in device memory of params->device_number (tensors of a training
 * step): same arguments and results, with device pointers and the caller's
 * stream (a hipStream_t, passed as void*; null = the null stream).
 * _create_dense_device validates and copies the arrays on the device (read
 * on `stream`, which is synchronised before it returns; INVALID_ARGUMENT
 * for pointers that are not device memory of that device);
 * _solution_device writes x, y, z (device, any may be null) with a kernel
 * on `stream` and returns without waiting; _adjoint_device runs the adjoint
 * kernels on `stream` with device dl_dx, u, v_rows, v_bounds and waits only
 * when `steps` (host) is given. Calls on one batch must not overlap. No
 * counterpart in the reference: the device-resident path of the
 * differentiable layer. */
HPRQPStatus hprqp_hip_batch_create_dense_device(size_t count, size_t n, size_t m, const double *Q, const double *c,
                                                const double *A, const double *AL, const double *AU, const double *l,
                                                const double *u, const HPRQPParams *params, void *stream,
                                                HPRQPHipBatch **out_batch);
HPRQPStatus hprqp_hip_batch_solution_device(const HPRQPHipBatch *batch, double *x, double *y, double *z, void *stream);
HPRQPStatus hprqp_hip_batch_adjoint_device(HPRQPHipBatch *batch, const double *dl_dx, double tol, int max_steps,
                                           double *u, double *v_rows, double *v_bounds, int *steps, void *stream);
/* Convergence of the batch's last adjoint (either form): per instance the
 * final relative residual ||K [u; v] - [g; 0]|| / ||[g; 0]|| of the MINRES
 * iterate (0 when g vanishes on the free variables); a value above the tol
 * passed to the adjoint means it stopped at max_steps or broke down. Read on
 * `stream` (the stream the device-buffer adjoint ran on) and waited for;
 * INVALID_ARGUMENT before any adjoint. No counterpart in the reference. */
HPRQPStatus hprqp_hip_batch_adjoint_residuals(const HPRQPHipBatch *batch, double *residuals, void *stream);

/* Copies `bytes` (a multiple of 16) between two fresh device buffers with a
 * plain streaming kernel and reports its time: the known-size kernel used to
 * calibrate HBM traffic counters. No counterpart in the reference. */
HPRQPStatus hprqp_hip_calibration_copy(size_t bytes, double *milliseconds);
/* Random-gather reference: mean time of one ELL product with 20 entries per
 * row over `nrow` rows, columns uniform over a `window`-element vector
 * (index and value loads coalesced). No reference counterpart: the ceiling
 * against which the gather-bound SpMV kernels are measured. */
HPRQPStatus hprqp_hip_calibration_gather(int nrow, int window, int reps, double *milliseconds);

/* Name of the device the library runs on, or "" when none is usable. */
const char *hprqp_hip_device_name(void);

#ifdef __cplusplus
}
#endif

#endif
