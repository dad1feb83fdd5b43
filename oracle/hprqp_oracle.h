/*
 * CPU oracle for the HPR-QP iteration: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU path (src/cpu/ *.c and
 * src/shared/{build,sparse,vector,solver,solver_backend}.c) plus the
 * second-order-cone step, which the reference only has on the GPU
 * (src/gpu/soc_{common,iteration,residuals}.cuh). Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may load it; the HIP product never
 * links or calls it.
 *
 * Interface: arrays in, opaque session out. A session owns a preprocessed,
 * scaled copy of the model plus the HPR state, and advances the reference's
 * solve loop one loop body at a time so that a caller can compare iterates.
 */
#ifndef HPRQP_ORACLE_H
#define HPRQP_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrSession OrSession;

typedef struct OrParams {
    double stoptol;
    size_t max_iter;
    double time_limit;
    size_t check_iter;
    int print_frequency;
    int use_scaling;
    /* > 0: use this lambda_max estimate instead of running the power
     * iteration (only the CPU-baseline timing uses this, to bound setup). */
    double lambda_A_given;
    double lambda_Q_given;
    /* equality-row (Gram) metric: 0 OFF, 1 AUTO (default), 2 REQUEST */
    int gram_mode;
} OrParams;

typedef struct OrResults {
    int status;
    size_t iter;
    size_t iter_4;
    size_t iter_6;
    double time;
    double time_4;
    double time_6;
    double primal_obj;
    double residuals;
    double gap;
    double err_rp;
    double err_rd;
    double dual_obj;
} OrResults;

/* Status codes match HPRQPStatus. */
enum { OR_OK = 0, OR_INVALID = 1, OR_ALLOC = 2, OR_NOT_IMPL = 3, OR_MAX_ITER = 4,
       OR_TIME_LIMIT = 5, OR_OPTIMAL = 6, OR_CONTINUE = 7 };

void or_params_defaults(OrParams *p);

/* Build (copy + preprocess, like build_from_QAbc[_soc]) then scale and set up
 * the workspace (sigma0, noC, diag-Q, power iteration). var/row cone offsets
 * may be NULL when the counts are 0. */
int or_session_create(size_t n, size_t m,
                      size_t q_nnz, const size_t *q_ptr, const size_t *q_ind, const double *q_val,
                      size_t a_nnz, const size_t *a_ptr, const size_t *a_ind, const double *a_val,
                      const double *c, const double *AL, const double *AU,
                      const double *l, const double *u, double obj_constant,
                      size_t num_var_cones, const size_t *var_offsets,
                      size_t num_row_cones, const size_t *row_offsets, const double *soc_rhs,
                      const OrParams *params, OrSession **out);
void or_session_destroy(OrSession *s);

/* Run up to `bodies` bodies of the solve loop. *finished becomes 1 when the
 * loop terminated (status in or_session_results). */
int or_session_advance(OrSession *s, size_t bodies, int *finished);
/* Run the loop to termination and fill results (original units in x/y/z/w). */
int or_session_solve(OrSession *s, OrResults *res, double *x, double *y, double *z, double *w);
int or_session_results(OrSession *s, OrResults *res, double *x, double *y, double *z, double *w);

/* Sizes after preprocessing. */
size_t or_session_n(const OrSession *s);
size_t or_session_m(const OrSession *s);
/* Read a named scaled-space workspace vector ("x", "x_bar", "w", "w_bar",
 * "y", "y_bar", "z_bar", "s", "ATy", "Qw", "last_x", ...) or a named model
 * array ("c", "l", "u", "AL", "AU", "row_norm", "col_norm"). Returns its
 * length, or 0 if unknown; copies min(len, cap) values. */
size_t or_session_read(const OrSession *s, const char *name, double *out, size_t cap);
/* Named scalars: sigma, lambda_max_A, lambda_max_Q, Q_is_diag, noC, iter,
 * inner, restart_times, norm_b_org, norm_c_org, kkt, err_rp, err_rd, gap,
 * pobj, dobj. NaN if unknown. */
double or_session_scalar(const OrSession *s, const char *name);
/* Scaled CSR of "Q", "A" or "AT": fills what is non-NULL, returns nnz. */
size_t or_session_matrix(const OrSession *s, const char *name, size_t *ptr, size_t *ind, double *val);

/* Preprocessing alone (symmetrise Q, drop empty/free rows) for checking the
 * builder: returns nnz of Q and A via out params and fills arrays when given
 * capacity. Returns status. */
int or_preprocess(size_t n, size_t m,
                  size_t q_nnz, const size_t *q_ptr, const size_t *q_ind, const double *q_val,
                  size_t a_nnz, const size_t *a_ptr, const size_t *a_ind, const double *a_val,
                  const double *AL, const double *AU,
                  size_t *q_nnz_out, size_t *q_ptr_out, size_t *q_ind_out, double *q_val_out,
                  size_t *m_out, size_t *a_nnz_out, size_t *a_ptr_out, size_t *a_ind_out, double *a_val_out);

/* y = M x with the sequential row sum of sparse.c; and the counting-sort
 * transpose. */
void or_csr_matvec(size_t nrow, const size_t *ptr, const size_t *ind, const double *val,
                   const double *x, double *y);
void or_csr_transpose(size_t nrow, size_t ncol, const size_t *ptr, const size_t *ind, const double *val,
                      size_t *tptr, size_t *tind, double *tval);

/* Closed-form unit checks: overwrite a named workspace vector, run one
 * iteration_step(inner) with the given to_check, evaluate residuals at `iter`
 * (out: pobj, dobj, err_rd, err_rp, gap, kkt), or the restart M-norm. */
int or_session_write(OrSession *s, const char *name, const double *in, size_t len);
int or_session_step(OrSession *s, size_t inner, int to_check);
int or_session_residuals(OrSession *s, size_t iter, double *out6);
double or_session_m_norm(OrSession *s);

/* Equality-row (Gram) metric: the quadratic 2||E'dy_E||^2 + 2 beta
 * ||dy_rest||^2 at the current dy (NaN when the metric is off), and the row
 * map (metric index or -1 per row; returns m, 0 when off). Vectors
 * "gram_G", "gram_r", "gram_step", "gram_b" and scalars "gram_active" (0 off,
 * 1 plain-QP metric, 2 row-cone prefix metric), "gram_rows", "gram_beta",
 * "gram_nnz", "gram_backward", "gram_mode" are readable by name. */
double or_session_gram_quadratic(OrSession *s);
size_t or_session_gram_map(const OrSession *s, long *out, size_t cap);

/* Lorentz-cone projection of one block (t = in[0], v = in[1:len]). */
void or_soc_project(double *out, const double *in, size_t len);

#ifdef __cplusplus
}
#endif

#endif
