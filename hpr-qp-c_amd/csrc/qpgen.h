/* Seeded synthetic QP generators (benchmark and test workloads). */
#ifndef HPRQP_QPGEN_H
#define HPRQP_QPGEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct QPGenProblem {
    size_t n, m, m_eq;
    size_t q_nnz, a_nnz;
    size_t *q_ptr, *q_ind;
    double *q_val;
    size_t *a_ptr, *a_ind;
    double *a_val;
    double *c, *AL, *AU, *l, *u, *x0;
    size_t num_var_cones;
    size_t *var_offsets;
    size_t num_row_cones;
    size_t *row_offsets;
    double *soc_rhs;
} QPGenProblem;

int qpgen_random_qp(uint64_t seed, size_t n, size_t m_eq, size_t m_in, size_t a_row_nnz, size_t heavy_rows,
                    size_t heavy_nnz, size_t b_rows, size_t b_row_nnz, double box, QPGenProblem *p);
int qpgen_soc_qp(uint64_t seed, size_t n_box, size_t cones, size_t cone_dim, size_t m, size_t a_row_nnz,
                 double box, QPGenProblem *p);
void qpgen_free(QPGenProblem *p);

#ifdef __cplusplus
}
#endif

#endif
