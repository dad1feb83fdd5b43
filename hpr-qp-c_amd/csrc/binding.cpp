// Opaque-handle binding API over the HPR-QP C API. Same entry points, argument
// meaning and status codes as the reference's src/shared/binding.c:1-346
// (declarations: include/hprqp_binding.h:12-79); own implementation.
#include <cstdlib>
#include <new>

#include "hprqp.h"
#include "hprqp_binding.h"

struct HPRQPBindingModel {
    HPRQPModel model;
};

struct HPRQPBindingParams {
    HPRQPParams params;
};

struct HPRQPBindingResults {
    HPRQPResults results;
    size_t n;
    size_t m;
    HPRQPStatus status;
};

namespace {

// binding.c:51-76: a non-owning CSR view; empty arrays point at a dummy.
HPRQPStatus csr_view(HPRQPMatrixCSR *mat, size_t nrow, size_t ncol, size_t nnz, const size_t *row_ptr,
                     const size_t *col_ind, const double *values, size_t *empty_index, double *empty_value)
{
    if ((nrow > 0 && row_ptr == nullptr) || (nnz > 0 && (col_ind == nullptr || values == nullptr)))
        return HPRQP_STATUS_INVALID_ARGUMENT;
    mat->nrow = nrow;
    mat->ncol = ncol;
    mat->nnz = nnz;
    mat->row_ptr = const_cast<size_t *>(row_ptr ? row_ptr : empty_index);
    mat->col_ind = const_cast<size_t *>(col_ind ? col_ind : empty_index);
    mat->values = const_cast<double *>(values ? values : empty_value);
    mat->owns_data = 0;
    return HPRQP_STATUS_OK;
}

}  // namespace

extern "C" {

const char *hprqp_status_name(HPRQPStatus status)
{
    switch (status) {
    case HPRQP_STATUS_OK: return "OK";
    case HPRQP_STATUS_INVALID_ARGUMENT: return "INVALID_ARGUMENT";
    case HPRQP_STATUS_ALLOCATION_FAILED: return "ALLOCATION_FAILED";
    case HPRQP_STATUS_NOT_IMPLEMENTED: return "NOT_IMPLEMENTED";
    case HPRQP_STATUS_MAX_ITER: return "MAX_ITER";
    case HPRQP_STATUS_TIME_LIMIT: return "TIME_LIMIT";
    case HPRQP_STATUS_OPTIMAL: return "OPTIMAL";
    case HPRQP_STATUS_CONTINUE: return "CONTINUE";
    default: return "UNKNOWN";
    }
}

int hprqp_gpu_backend_available(void) { return gpu_backend_available(); }

HPRQPStatus hprqp_model_create(HPRQPBindingModel **out_model)
{
    if (!out_model) return HPRQP_STATUS_INVALID_ARGUMENT;
    auto *m = static_cast<HPRQPBindingModel *>(std::malloc(sizeof(HPRQPBindingModel)));
    if (!m) return HPRQP_STATUS_ALLOCATION_FAILED;
    model_init(&m->model);
    *out_model = m;
    return HPRQP_STATUS_OK;
}

void hprqp_model_destroy(HPRQPBindingModel *model)
{
    if (!model) return;
    model_free(&model->model);
    std::free(model);
}

HPRQPStatus hprqp_model_read_from_mps(HPRQPBindingModel *model, const char *mps_path)
{
    if (!model || !mps_path) return HPRQP_STATUS_INVALID_ARGUMENT;
    return build_from_mps(mps_path, &model->model);
}

HPRQPStatus hprqp_model_build_from_qabc_csr(HPRQPBindingModel *model, size_t q_nrow, size_t q_ncol, size_t q_nnz,
                                            const size_t *q_row_ptr, const size_t *q_col_ind,
                                            const double *q_values, size_t a_nrow, size_t a_ncol, size_t a_nnz,
                                            const size_t *a_row_ptr, const size_t *a_col_ind,
                                            const double *a_values, const double *c, const double *AL,
                                            const double *AU, const double *l, const double *u,
                                            double obj_constant)
{
    if (!model || !c || !l || !u) return HPRQP_STATUS_INVALID_ARGUMENT;
    HPRQPMatrixCSR q, a;
    size_t q_empty_index = 0, a_empty_index = 0;
    double q_empty_value = 0.0, a_empty_value = 0.0;
    HPRQPStatus st = csr_view(&q, q_nrow, q_ncol, q_nnz, q_row_ptr, q_col_ind, q_values, &q_empty_index,
                              &q_empty_value);
    if (st != HPRQP_STATUS_OK) return st;
    st = csr_view(&a, a_nrow, a_ncol, a_nnz, a_row_ptr, a_col_ind, a_values, &a_empty_index, &a_empty_value);
    if (st != HPRQP_STATUS_OK) return st;
    return build_from_QAbc(&q, &a, c, AL, AU, l, u, obj_constant, &model->model);
}

size_t hprqp_model_num_variables(const HPRQPBindingModel *model) { return model ? model->model.Q.ncol : 0; }
size_t hprqp_model_num_constraints(const HPRQPBindingModel *model) { return model ? model->model.A.nrow : 0; }

HPRQPStatus hprqp_params_create(HPRQPBindingParams **out_params)
{
    if (!out_params) return HPRQP_STATUS_INVALID_ARGUMENT;
    auto *p = static_cast<HPRQPBindingParams *>(std::malloc(sizeof(HPRQPBindingParams)));
    if (!p) return HPRQP_STATUS_ALLOCATION_FAILED;
    params_set_defaults(&p->params);
    *out_params = p;
    return HPRQP_STATUS_OK;
}

void hprqp_params_destroy(HPRQPBindingParams *params) { std::free(params); }

HPRQPStatus hprqp_params_set_defaults(HPRQPBindingParams *params)
{
    if (!params) return HPRQP_STATUS_INVALID_ARGUMENT;
    params_set_defaults(&params->params);
    return HPRQP_STATUS_OK;
}

#define HPRQP_SETTER(NAME, TYPE)                                                   \
    HPRQPStatus hprqp_params_set_##NAME(HPRQPBindingParams *params, TYPE value)    \
    {                                                                              \
        if (!params) return HPRQP_STATUS_INVALID_ARGUMENT;                         \
        params->params.NAME = value;                                               \
        return HPRQP_STATUS_OK;                                                    \
    }
HPRQP_SETTER(stoptol, double)
HPRQP_SETTER(max_iter, size_t)
HPRQP_SETTER(use_gpu, int)
HPRQP_SETTER(device_number, int)
HPRQP_SETTER(time_limit, double)
HPRQP_SETTER(check_iter, size_t)
HPRQP_SETTER(print_frequency, int)
HPRQP_SETTER(use_scaling, int)
HPRQP_SETTER(verbose, int)
HPRQP_SETTER(spmv_mode_q, HPRQPSpmvMode)
HPRQP_SETTER(spmv_mode_a, HPRQPSpmvMode)
#undef HPRQP_SETTER

HPRQPStatus hprqp_results_create(HPRQPBindingResults **out_results)
{
    if (!out_results) return HPRQP_STATUS_INVALID_ARGUMENT;
    auto *r = static_cast<HPRQPBindingResults *>(std::malloc(sizeof(HPRQPBindingResults)));
    if (!r) return HPRQP_STATUS_ALLOCATION_FAILED;
    results_init(&r->results);
    r->n = r->m = 0;
    r->status = HPRQP_STATUS_OK;
    *out_results = r;
    return HPRQP_STATUS_OK;
}

void hprqp_results_destroy(HPRQPBindingResults *results)
{
    if (!results) return;
    results_free(&results->results);
    std::free(results);
}

// binding.c:297-310
HPRQPStatus hprqp_optimize(const HPRQPBindingModel *model, const HPRQPBindingParams *params,
                           HPRQPBindingResults *results)
{
    if (!model || !params || !results) return HPRQP_STATUS_INVALID_ARGUMENT;
    results_free(&results->results);
    results->n = model->model.Q.ncol;
    results->m = model->model.A.nrow;
    results->status = solve(&model->model, &params->params, &results->results);
    return results->status;
}

HPRQPStatus hprqp_results_status(const HPRQPBindingResults *r) { return r ? r->status : HPRQP_STATUS_INVALID_ARGUMENT; }
size_t hprqp_results_num_variables(const HPRQPBindingResults *r) { return r ? r->n : 0; }
size_t hprqp_results_num_constraints(const HPRQPBindingResults *r) { return r ? r->m : 0; }
size_t hprqp_results_iter(const HPRQPBindingResults *r) { return r ? r->results.iter : 0; }
size_t hprqp_results_iter_4(const HPRQPBindingResults *r) { return r ? r->results.iter_4 : 0; }
size_t hprqp_results_iter_6(const HPRQPBindingResults *r) { return r ? r->results.iter_6 : 0; }
double hprqp_results_time(const HPRQPBindingResults *r) { return r ? r->results.time : 0.0; }
double hprqp_results_time_4(const HPRQPBindingResults *r) { return r ? r->results.time_4 : 0.0; }
double hprqp_results_time_6(const HPRQPBindingResults *r) { return r ? r->results.time_6 : 0.0; }
double hprqp_results_objective(const HPRQPBindingResults *r) { return r ? r->results.primal_obj : 0.0; }
double hprqp_results_residual(const HPRQPBindingResults *r) { return r ? r->results.residuals : 0.0; }
double hprqp_results_gap(const HPRQPBindingResults *r) { return r ? r->results.gap : 0.0; }
const double *hprqp_results_x_data(const HPRQPBindingResults *r) { return r ? r->results.x : nullptr; }
const double *hprqp_results_y_data(const HPRQPBindingResults *r) { return r ? r->results.y : nullptr; }
const double *hprqp_results_z_data(const HPRQPBindingResults *r) { return r ? r->results.z : nullptr; }
const double *hprqp_results_w_data(const HPRQPBindingResults *r) { return r ? r->results.w : nullptr; }

}  // extern "C"
