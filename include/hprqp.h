/*
 * The HPR-QP C API (reference include/hprqp.h:13-39), implemented by
 * libhprqp_hip.so on MI355X. Each entry point replaces the reference symbol of
 * the same name:
 *   model_init/free/copy           src/shared/build.c:402-502
 *   build_from_QAbc[_soc]          src/shared/build.c:602-621
 *   build_from_mps                 src/shared/mps_qp.cpp:206-326 with
 *                                  src/shared/mps_reader.cpp
 *   build_from_QAbc_qc_factors,    src/shared/qcqp.c:16-99 (declared in
 *   build_from_QAbc_qc_diagonal    shared/qcqp.h, included below)
 *   params_set_defaults            src/shared/solver.c:133-157
 *   results_init/free              src/shared/solver.c:159-177
 *   print_model_info/params_info   src/shared/solver_backend.c:240-271
 *   solve_gpu                      src/gpu/solver.cu:127-231
 *   solve                          src/shared/solver.c:245-260 (use_gpu=0
 *                                  returns HPRQP_STATUS_NOT_IMPLEMENTED: this
 *                                  library carries no CPU solver)
 */
#ifndef HPRQP_H
#define HPRQP_H

#ifdef __cplusplus
extern "C" {
#endif

#include "gpu/gpu.h"
#include "shared/qcqp.h"
#include "shared/sparse.h"
#include "shared/types.h"

void model_init(HPRQPModel *model);
void model_free(HPRQPModel *model);
HPRQPStatus model_copy(HPRQPModel *dest, const HPRQPModel *src);
HPRQPStatus build_from_QAbc(const HPRQPMatrixCSR *Q,
                            const HPRQPMatrixCSR *A,
                            const double *c,
                            const double *AL,
                            const double *AU,
                            const double *l,
                            const double *u,
                            double obj_constant,
                            HPRQPModel *model);
HPRQPStatus build_from_QAbc_soc(const HPRQPMatrixCSR *Q,
                                const HPRQPMatrixCSR *A,
                                const double *c, const double *AL, const double *AU,
                                const double *l, const double *u, double obj_constant,
                                const HPRQPSOCSpec *soc, HPRQPModel *model);
HPRQPStatus build_from_mps(const char *mps_path, HPRQPModel *model);

void params_set_defaults(HPRQPParams *params);
void results_init(HPRQPResults *results);
void results_free(HPRQPResults *results);
void print_model_info(const HPRQPModel *model);
void print_params_info(const HPRQPParams *params);
HPRQPStatus solve_gpu(const HPRQPModel *model, const HPRQPParams *params, HPRQPResults *results);
HPRQPStatus solve(const HPRQPModel *model, const HPRQPParams *params, HPRQPResults *results);

#ifdef __cplusplus
}
#endif

#endif
