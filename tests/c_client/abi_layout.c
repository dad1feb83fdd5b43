/* Prints the size, alignment and field offsets of the ABI structs of
 * shared/types.h and the values of its enums, one "name value" line each.
 * Compiled once against this repository's include/ and once against the
 * reference's include/ (include/shared/types.h:10-214): the two outputs must
 * be identical for a client built against one to run with the other. */
#include <stddef.h>
#include <stdio.h>

#include "shared/types.h"

#define SIZE(T) printf("sizeof(%s) %zu\nalignof(%s) %zu\n", #T, sizeof(T), #T, _Alignof(T))
#define OFF(T, f) printf("offsetof(%s,%s) %zu\n", #T, #f, offsetof(T, f))
#define VAL(e) printf("%s %d\n", #e, (int)(e))

int main(void)
{
    SIZE(HPRQPMatrixCSR);
    OFF(HPRQPMatrixCSR, nrow); OFF(HPRQPMatrixCSR, ncol); OFF(HPRQPMatrixCSR, nnz);
    OFF(HPRQPMatrixCSR, row_ptr); OFF(HPRQPMatrixCSR, col_ind); OFF(HPRQPMatrixCSR, values);
    OFF(HPRQPMatrixCSR, owns_data);

    SIZE(HPRQPSOCSpec);
    OFF(HPRQPSOCSpec, num_var_cones); OFF(HPRQPSOCSpec, var_offsets); OFF(HPRQPSOCSpec, num_row_cones);
    OFF(HPRQPSOCSpec, row_offsets); OFF(HPRQPSOCSpec, rhs);

    SIZE(HPRQPModel);
    OFF(HPRQPModel, Q); OFF(HPRQPModel, A); OFF(HPRQPModel, AT); OFF(HPRQPModel, c); OFF(HPRQPModel, AL);
    OFF(HPRQPModel, AU); OFF(HPRQPModel, l); OFF(HPRQPModel, u); OFF(HPRQPModel, obj_constant);
    OFF(HPRQPModel, owns_data); OFF(HPRQPModel, num_var_cones); OFF(HPRQPModel, var_cone_offsets);
    OFF(HPRQPModel, num_row_cones); OFF(HPRQPModel, row_cone_offsets); OFF(HPRQPModel, soc_rhs);

    SIZE(HPRQPScaling);
    OFF(HPRQPScaling, row_norm); OFF(HPRQPScaling, col_norm); OFF(HPRQPScaling, l_org); OFF(HPRQPScaling, u_org);
    OFF(HPRQPScaling, b_scale); OFF(HPRQPScaling, c_scale); OFF(HPRQPScaling, norm_b); OFF(HPRQPScaling, norm_c);
    OFF(HPRQPScaling, norm_b_org); OFF(HPRQPScaling, norm_c_org);

    SIZE(HPRQPParams);
    OFF(HPRQPParams, stoptol); OFF(HPRQPParams, max_iter); OFF(HPRQPParams, use_gpu);
    OFF(HPRQPParams, device_number); OFF(HPRQPParams, time_limit); OFF(HPRQPParams, check_iter);
    OFF(HPRQPParams, print_frequency); OFF(HPRQPParams, use_scaling); OFF(HPRQPParams, verbose);
    OFF(HPRQPParams, spmv_mode_q); OFF(HPRQPParams, spmv_mode_a); OFF(HPRQPParams, gram_mode);
    OFF(HPRQPParams, polishing); OFF(HPRQPParams, polish_python); OFF(HPRQPParams, polish_worker);
    OFF(HPRQPParams, polish_directory); OFF(HPRQPParams, polish_keep_files); OFF(HPRQPParams, polish_max_attempts);
    OFF(HPRQPParams, polish_episode_seconds); OFF(HPRQPParams, polish_total_seconds);
    OFF(HPRQPParams, polish_memory_bytes);

    SIZE(HPRQPWorkspace);
    OFF(HPRQPWorkspace, m); OFF(HPRQPWorkspace, n); OFF(HPRQPWorkspace, sigma); OFF(HPRQPWorkspace, lambda_max_A);
    OFF(HPRQPWorkspace, lambda_max_Q); OFF(HPRQPWorkspace, Q_is_diag); OFF(HPRQPWorkspace, noC);
    OFF(HPRQPWorkspace, to_check); OFF(HPRQPWorkspace, Q); OFF(HPRQPWorkspace, c); OFF(HPRQPWorkspace, x);
    OFF(HPRQPWorkspace, y); OFF(HPRQPWorkspace, z_bar); OFF(HPRQPWorkspace, fact_M); OFF(HPRQPWorkspace, last_ATy);

    SIZE(HPRQPResiduals);
    OFF(HPRQPResiduals, is_updated); OFF(HPRQPResiduals, err_Rp_org_bar); OFF(HPRQPResiduals, err_Rd_org_bar);
    OFF(HPRQPResiduals, KKTx_and_gap_org_bar); OFF(HPRQPResiduals, primal_obj_bar); OFF(HPRQPResiduals, rel_gap_bar);
    OFF(HPRQPResiduals, dual_obj_bar);

    SIZE(HPRQPRestart);
    OFF(HPRQPRestart, restart_flag); OFF(HPRQPRestart, first_restart); OFF(HPRQPRestart, last_gap);
    OFF(HPRQPRestart, current_gap); OFF(HPRQPRestart, save_gap); OFF(HPRQPRestart, inner); OFF(HPRQPRestart, step);
    OFF(HPRQPRestart, sufficient); OFF(HPRQPRestart, necessary); OFF(HPRQPRestart, long_restart);
    OFF(HPRQPRestart, ratio); OFF(HPRQPRestart, times); OFF(HPRQPRestart, weighted_norm);
    OFF(HPRQPRestart, best_gap); OFF(HPRQPRestart, best_sigma);

    SIZE(HPRQPResults);
    OFF(HPRQPResults, iter); OFF(HPRQPResults, iter_4); OFF(HPRQPResults, iter_6); OFF(HPRQPResults, time);
    OFF(HPRQPResults, time_4); OFF(HPRQPResults, time_6); OFF(HPRQPResults, primal_obj);
    OFF(HPRQPResults, residuals); OFF(HPRQPResults, gap); OFF(HPRQPResults, status); OFF(HPRQPResults, x);
    OFF(HPRQPResults, y); OFF(HPRQPResults, z); OFF(HPRQPResults, w); OFF(HPRQPResults, polish_attempts);
    OFF(HPRQPResults, polish_accepted); OFF(HPRQPResults, gram_active); OFF(HPRQPResults, total_time);

    VAL(HPRQP_STATUS_OK); VAL(HPRQP_STATUS_INVALID_ARGUMENT); VAL(HPRQP_STATUS_ALLOCATION_FAILED);
    VAL(HPRQP_STATUS_NOT_IMPLEMENTED); VAL(HPRQP_STATUS_MAX_ITER); VAL(HPRQP_STATUS_TIME_LIMIT);
    VAL(HPRQP_STATUS_OPTIMAL); VAL(HPRQP_STATUS_CONTINUE);
    VAL(HPRQP_SPMV_MODE_AUTO); VAL(HPRQP_SPMV_MODE_CUSPARSE); VAL(HPRQP_SPMV_MODE_CUSTOMIZED);
    VAL(HPRQP_GRAM_OFF); VAL(HPRQP_GRAM_AUTO); VAL(HPRQP_GRAM_REQUEST);
    return 0;
}
