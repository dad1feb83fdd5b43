/* Convex quadratic constraints lifted to rotated second-order cones; same
 * declarations as the reference's include/shared/qcqp.h:7-22
 * (implementations: src/shared/qcqp.c:16-99).
 *
 * Constraint i reads  h_i(x) = 0.5*||R_i x||^2 + C_i x + constant_i <= 0,
 * where R_i is the block of rows offsets[i] .. offsets[i+1]-1 of R. Each
 * constraint becomes one row cone of rows(R_i) + 2 rows appended after A's
 * rows; no variable is added and R_i'R_i is never formed. */
#ifndef HPRQP_QCQP_H
#define HPRQP_QCQP_H

#include "shared/types.h"

#ifdef __cplusplus
extern "C" {
#endif

HPRQPStatus build_from_QAbc_qc_factors(const HPRQPMatrixCSR *Q, const HPRQPMatrixCSR *A, const double *c,
                                       const double *AL, const double *AU, const double *l, const double *u,
                                       double obj_constant, const HPRQPMatrixCSR *C, const double *constant,
                                       const HPRQPMatrixCSR *R, const size_t *offsets, HPRQPModel *model);

/* Diagonal form: h_i(x) = 0.5*sum_j D_ij x_j^2 + C_i x + constant_i <= 0 with
 * finite D >= 0; every positive D_ij becomes the factor row sqrt(D_ij) e_j. */
HPRQPStatus build_from_QAbc_qc_diagonal(const HPRQPMatrixCSR *Q, const HPRQPMatrixCSR *A, const double *c,
                                        const double *AL, const double *AU, const double *l, const double *u,
                                        double obj_constant, const HPRQPMatrixCSR *C, const double *constant,
                                        const HPRQPMatrixCSR *D, HPRQPModel *model);

#ifdef __cplusplus
}
#endif

#endif
