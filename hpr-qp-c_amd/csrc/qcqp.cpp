// Convex quadratic constraints as row cones (src/shared/qcqp.c:16-99).
//
// h(x) = 0.5*||Rx||^2 + Cx + k <= 0 holds exactly when
//     t = 1 - (Cx + k),  v = -1 - (Cx + k),  tail = sqrt(2)*Rx
// satisfies t >= ||(v, tail)||: t^2 - v^2 - ||tail||^2 = -4h and
// t - v = 2 > 0. As rows of the lifted model, both head rows carry -C with
// cone right-hand sides k - 1 and k + 1, and the tail rows carry sqrt(2)*R
// (right-hand side 0). The cone rows are appended after A's rows, one cone
// per constraint, with (-inf, inf) row bounds; build_from_QAbc_soc then
// preprocesses the lifted model. The coefficients are formed as the
// reference forms them (sqrt(2) times the stored factor; the diagonal form
// stores sqrt(D)), so both builds give bit-identical models.
#include <cmath>
#include <cstdint>
#include <new>
#include <vector>

#include "hprqp.h"

namespace {

bool finite_csr(const HPRQPMatrixCSR *M)
{
    if (!matrix_csr_is_valid(M)) return false;
    for (size_t k = 0; k < M->nnz; ++k)
        if (!std::isfinite(M->values[k])) return false;
    return true;
}

}  // namespace

extern "C" {

HPRQPStatus build_from_QAbc_qc_factors(const HPRQPMatrixCSR *Q, const HPRQPMatrixCSR *A, const double *c,
                                       const double *AL, const double *AU, const double *l, const double *u,
                                       double obj_constant, const HPRQPMatrixCSR *C, const double *constant,
                                       const HPRQPMatrixCSR *R, const size_t *offsets, HPRQPModel *model)
{
    if (!Q || !A || !model || !finite_csr(C) || !finite_csr(R) || !matrix_csr_is_valid(A)) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (C->ncol != Q->ncol || R->ncol != Q->ncol || A->ncol != Q->ncol) return HPRQP_STATUS_INVALID_ARGUMENT;
    const size_t nq = C->nrow, m = A->nrow;
    if (nq == 0) return build_from_QAbc(Q, A, c, AL, AU, l, u, obj_constant, model);
    if (!constant || !offsets || offsets[0] != 0 || offsets[nq] != R->nrow || (m && (!AL || !AU)))
        return HPRQP_STATUS_INVALID_ARGUMENT;
    for (size_t i = 0; i < nq; ++i)
        if (!std::isfinite(constant[i]) || offsets[i] > offsets[i + 1] || offsets[i + 1] > R->nrow)
            return HPRQP_STATUS_INVALID_ARGUMENT;
    // sizes of the lifted model must fit size_t
    if (R->nrow > SIZE_MAX - m || nq > (SIZE_MAX - m - R->nrow) / 2 || R->nnz > SIZE_MAX - A->nnz ||
        C->nnz > (SIZE_MAX - A->nnz - R->nnz) / 2)
        return HPRQP_STATUS_INVALID_ARGUMENT;
    const size_t rows = m + 2 * nq + R->nrow, nnz = A->nnz + 2 * C->nnz + R->nnz;
    if (rows > SIZE_MAX / sizeof(double) || nnz > SIZE_MAX / sizeof(double)) return HPRQP_STATUS_INVALID_ARGUMENT;
    try {
        std::vector<size_t> rp, ci, cones;
        std::vector<double> val, lo, hi, rhs(rows, 0.0);
        rp.reserve(rows + 1);
        ci.reserve(nnz + 1);  // never empty: a valid CSR has non-null arrays
        val.reserve(nnz + 1);
        cones.reserve(nq + 1);
        for (size_t i = 0; i < m; ++i) {
            rp.push_back(ci.size());
            for (size_t k = A->row_ptr[i]; k < A->row_ptr[i + 1]; ++k) {
                ci.push_back(A->col_ind[k]);
                val.push_back(A->values[k]);
            }
        }
        lo.assign(AL, AL + m);
        hi.assign(AU, AU + m);
        const double root2 = std::sqrt(2.0);
        for (size_t q = 0; q < nq; ++q) {
            cones.push_back(rp.size());
            for (const double shift : {-1.0, 1.0}) {  // head rows t, then v
                rhs[rp.size()] = constant[q] + shift;
                rp.push_back(ci.size());
                for (size_t k = C->row_ptr[q]; k < C->row_ptr[q + 1]; ++k) {
                    ci.push_back(C->col_ind[k]);
                    val.push_back(-C->values[k]);
                }
            }
            for (size_t r = offsets[q]; r < offsets[q + 1]; ++r) {
                rp.push_back(ci.size());
                for (size_t k = R->row_ptr[r]; k < R->row_ptr[r + 1]; ++k) {
                    ci.push_back(R->col_ind[k]);
                    val.push_back(root2 * R->values[k]);
                }
            }
        }
        cones.push_back(rp.size());
        rp.push_back(ci.size());
        lo.resize(rows, -INFINITY);
        hi.resize(rows, INFINITY);
        const HPRQPMatrixCSR lifted{rows, Q->ncol, ci.size(), rp.data(), ci.data(), val.data(), 0};
        const HPRQPSOCSpec spec{0, nullptr, nq, cones.data(), rhs.data()};
        return build_from_QAbc_soc(Q, &lifted, c, lo.data(), hi.data(), l, u, obj_constant, &spec, model);
    } catch (const std::bad_alloc &) {
        return HPRQP_STATUS_ALLOCATION_FAILED;
    }
}

HPRQPStatus build_from_QAbc_qc_diagonal(const HPRQPMatrixCSR *Q, const HPRQPMatrixCSR *A, const double *c,
                                        const double *AL, const double *AU, const double *l, const double *u,
                                        double obj_constant, const HPRQPMatrixCSR *C, const double *constant,
                                        const HPRQPMatrixCSR *D, HPRQPModel *model)
{
    if (!Q || !C || !finite_csr(D) || C->nrow != D->nrow || C->ncol != D->ncol || D->ncol != Q->ncol)
        return HPRQP_STATUS_INVALID_ARGUMENT;
    if (D->nrow >= SIZE_MAX / sizeof(size_t)) return HPRQP_STATUS_INVALID_ARGUMENT;
    for (size_t k = 0; k < D->nnz; ++k)
        if (D->values[k] < 0.0) return HPRQP_STATUS_INVALID_ARGUMENT;
    try {
        // one factor row sqrt(D_ij) e_j per positive entry, grouped by constraint
        std::vector<size_t> rp{0}, ci, offsets;
        std::vector<double> val;
        offsets.reserve(D->nrow + 1);
        ci.reserve(D->nnz + 1);  // never empty: a valid CSR has non-null arrays
        val.reserve(D->nnz + 1);
        for (size_t i = 0; i < D->nrow; ++i) {
            offsets.push_back(ci.size());
            for (size_t k = D->row_ptr[i]; k < D->row_ptr[i + 1]; ++k)
                if (D->values[k] > 0.0) {
                    ci.push_back(D->col_ind[k]);
                    val.push_back(std::sqrt(D->values[k]));
                    rp.push_back(ci.size());
                }
        }
        offsets.push_back(ci.size());
        const HPRQPMatrixCSR R{ci.size(), D->ncol, ci.size(), rp.data(), ci.data(), val.data(), 0};
        return build_from_QAbc_qc_factors(Q, A, c, AL, AU, l, u, obj_constant, C, constant, &R, offsets.data(),
                                          model);
    } catch (const std::bad_alloc &) {
        return HPRQP_STATUS_ALLOCATION_FAILED;
    }
}

}  // extern "C"
