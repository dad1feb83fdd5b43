/*
 * Public types of the HPR-QP C API. The layouts below are ABI: they must match
 * the reference's include/shared/types.h:10-214 field for field so that code
 * compiled against the reference links against this library unchanged.
 */
#ifndef HPRQP_TYPES_H
#define HPRQP_TYPES_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* types.h:10-19 */
typedef enum HPRQPStatus {
    HPRQP_STATUS_OK = 0,
    HPRQP_STATUS_INVALID_ARGUMENT = 1,
    HPRQP_STATUS_ALLOCATION_FAILED = 2,
    HPRQP_STATUS_NOT_IMPLEMENTED = 3,
    HPRQP_STATUS_MAX_ITER = 4,
    HPRQP_STATUS_TIME_LIMIT = 5,
    HPRQP_STATUS_OPTIMAL = 6,
    HPRQP_STATUS_CONTINUE = 7
} HPRQPStatus;

/* types.h:21-25. On this backend every mode selects the same hand-written
 * row-blocked SpMV; the field is kept for source compatibility. */
typedef enum HPRQPSpmvMode {
    HPRQP_SPMV_MODE_AUTO = 0,
    HPRQP_SPMV_MODE_CUSPARSE = 1,
    HPRQP_SPMV_MODE_CUSTOMIZED = 2
} HPRQPSpmvMode;

/* types.h:27-35: host CSR with size_t indices. */
typedef struct HPRQPMatrixCSR {
    size_t nrow;
    size_t ncol;
    size_t nnz;
    size_t *row_ptr;
    size_t *col_ind;
    double *values;
    int owns_data;
} HPRQPMatrixCSR;

/* types.h:37-48: standard Lorentz cones as contiguous suffix blocks of the
 * variables (var_offsets) and rows (row_offsets), count+1 offsets each,
 * ending at n or m; rhs has m entries, zero on the linear rows. */
typedef struct HPRQPSOCSpec {
    size_t num_var_cones;
    const size_t *var_offsets;
    size_t num_row_cones;
    const size_t *row_offsets;
    const double *rhs;
} HPRQPSOCSpec;

/* types.h:50-66 */
typedef struct HPRQPModel {
    HPRQPMatrixCSR Q;
    HPRQPMatrixCSR A;
    HPRQPMatrixCSR AT;
    double *c;
    double *AL;
    double *AU;
    double *l;
    double *u;
    double obj_constant;
    int owns_data;
    size_t num_var_cones;
    size_t *var_cone_offsets;
    size_t num_row_cones;
    size_t *row_cone_offsets;
    double *soc_rhs;
} HPRQPModel;

/* types.h:68-79 */
typedef struct HPRQPScaling {
    double *row_norm;
    double *col_norm;
    double *l_org;
    double *u_org;
    double b_scale;
    double c_scale;
    double norm_b;
    double norm_c;
    double norm_b_org;
    double norm_c_org;
} HPRQPScaling;

/* types.h:81-86. Only OFF is implemented here; AUTO resolves to OFF. */
typedef enum HPRQPGramMode {
    HPRQP_GRAM_OFF = 0,
    HPRQP_GRAM_AUTO = 1,
    HPRQP_GRAM_REQUEST = 2
} HPRQPGramMode;

/* types.h:88-110 */
typedef struct HPRQPParams {
    double stoptol;
    size_t max_iter;
    int use_gpu;
    int device_number;
    double time_limit;
    size_t check_iter;
    int print_frequency;
    int use_scaling;
    int verbose;
    HPRQPSpmvMode spmv_mode_q;
    HPRQPSpmvMode spmv_mode_a;
    HPRQPGramMode gram_mode;
    int polishing;
    const char *polish_python;
    const char *polish_worker;
    const char *polish_directory;
    int polish_keep_files;
    size_t polish_max_attempts;
    double polish_episode_seconds;
    double polish_total_seconds;
    size_t polish_memory_bytes;
} HPRQPParams;

/* types.h:112-164: host mirror of the iteration state (the HIP backend keeps
 * its own device-resident copy; this struct is part of the header surface). */
typedef struct HPRQPWorkspace {
    size_t m;
    size_t n;
    double sigma;
    double lambda_max_A;
    double lambda_max_Q;
    int Q_is_diag;
    int noC;
    int to_check;
    const HPRQPMatrixCSR *Q;
    const HPRQPMatrixCSR *A;
    const HPRQPMatrixCSR *AT;
    const double *c;
    const double *AL;
    const double *AU;
    const double *l;
    const double *u;
    double *x;
    double *x_bar;
    double *x_hat;
    double *w;
    double *w_bar;
    double *w_hat;
    double *y;
    double *y_bar;
    double *y_hat;
    double *Ax;
    double *ATy;
    double *ATy_bar;
    double *ATdy;
    double *Qw;
    double *Qw_bar;
    double *Qx;
    double *dQw;
    double *QATdy;
    double *dx;
    double *dy;
    double *dw;
    double *Rp;
    double *Rd;
    double *s;
    double *z_bar;
    double *tempv;
    double *diag_Q;
    double *fact1;
    double *fact2;
    double *fact;
    double *fact_M;
    double *last_x;
    double *last_y;
    double *last_Qw;
    double *last_w;
    double *last_ATy;
} HPRQPWorkspace;

/* types.h:166-174 */
typedef struct HPRQPResiduals {
    int is_updated;
    double err_Rp_org_bar;
    double err_Rd_org_bar;
    double KKTx_and_gap_org_bar;
    double primal_obj_bar;
    double rel_gap_bar;
    double dual_obj_bar;
} HPRQPResiduals;

/* types.h:176-193 */
typedef struct HPRQPRestart {
    int restart_flag;
    int first_restart;
    double last_gap;
    double current_gap;
    double save_gap;
    size_t inner;
    size_t step;
    size_t sufficient;
    size_t necessary;
    size_t long_restart;
    size_t ratio;
    size_t times;
    double weighted_norm;
    double best_gap;
    double best_sigma;
} HPRQPRestart;

/* types.h:195-214 */
typedef struct HPRQPResults {
    size_t iter;
    size_t iter_4;
    size_t iter_6;
    double time;
    double time_4;
    double time_6;
    double primal_obj;
    double residuals;
    double gap;
    int status;
    double *x;
    double *y;
    double *z;
    double *w;
    size_t polish_attempts;
    int polish_accepted;
    int gram_active;
    double total_time;
} HPRQPResults;

#ifdef __cplusplus
}
#endif

#endif
