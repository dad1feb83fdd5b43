/* A plain C caller of libhprqp_hip.so, compiled against this repository's
 * include/ only, the way a client of the reference's C API is (the cases of
 * the reference's python/tests/test_hprqp.py:12-60, which drive
 * src/shared/binding.c, plus the model API of include/hprqp.h that
 * examples/solve_qp.c uses). Each case prints one line of key=value pairs;
 * the test that runs it parses them.
 *
 * usage: client <path to simple_qp.mps>
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "hprqp.h"
#include "hprqp_binding.h"

static void print_binding(const char *name, HPRQPStatus rc, const HPRQPBindingResults *r)
{
    const size_t n = hprqp_results_num_variables(r), m = hprqp_results_num_constraints(r);
    const double *x = hprqp_results_x_data(r);
    printf("%s rc=%s status=%s n=%zu m=%zu iter=%zu objective=%.17g residual=%.17g gap=%.17g", name,
           hprqp_status_name(rc), hprqp_status_name(hprqp_results_status(r)), n, m, hprqp_results_iter(r),
           hprqp_results_objective(r), hprqp_results_residual(r), hprqp_results_gap(r));
    for (size_t j = 0; j < n && x; ++j) printf(" x%zu=%.17g", j, x[j]);
    printf("\n");
}

/* python/tests/test_hprqp.py:27-43: min x'x - 4x0 - 2x1, no rows, free
 * variables; x = (2, 1), objective -5, y empty. */
static int binding_unconstrained(void)
{
    const size_t q_ptr[3] = {0, 1, 2}, q_ind[2] = {0, 1};
    const double q_val[2] = {2.0, 2.0}, c[2] = {-4.0, -2.0};
    const size_t a_ptr[1] = {0};
    const double l[2] = {-INFINITY, -INFINITY}, u[2] = {INFINITY, INFINITY};
    HPRQPBindingModel *model = NULL;
    HPRQPBindingParams *params = NULL;
    HPRQPBindingResults *res = NULL;
    if (hprqp_model_create(&model) || hprqp_params_create(&params) || hprqp_results_create(&res)) return 1;
    HPRQPStatus st = hprqp_model_build_from_qabc_csr(model, 2, 2, 2, q_ptr, q_ind, q_val, 0, 2, 0, a_ptr, NULL, NULL,
                                                     c, NULL, NULL, l, u, 0.0);
    if (st != HPRQP_STATUS_OK) return 2;
    hprqp_params_set_verbose(params, 0);
    hprqp_params_set_stoptol(params, 1e-8);
    hprqp_params_set_use_gpu(params, 1);
    hprqp_params_set_device_number(params, 0);
    st = hprqp_optimize(model, params, res);
    print_binding("binding_unconstrained", st, res);
    hprqp_results_destroy(res);
    hprqp_params_destroy(params);
    hprqp_model_destroy(model);
    return 0;
}

/* python/tests/test_hprqp.py:16-25: simple_qp.mps read through the binding,
 * solved with max_iter 0: MAX_ITER, two variables, two rows. */
static int binding_mps_max_iter(const char *path)
{
    HPRQPBindingModel *model = NULL;
    HPRQPBindingParams *params = NULL;
    HPRQPBindingResults *res = NULL;
    if (hprqp_model_create(&model) || hprqp_params_create(&params) || hprqp_results_create(&res)) return 1;
    HPRQPStatus st = hprqp_model_read_from_mps(model, path);
    printf("binding_mps_read rc=%s n=%zu m=%zu\n", hprqp_status_name(st), hprqp_model_num_variables(model),
           hprqp_model_num_constraints(model));
    if (st != HPRQP_STATUS_OK) return 2;
    hprqp_params_set_defaults(params);
    hprqp_params_set_verbose(params, 0);
    hprqp_params_set_max_iter(params, 0);
    st = hprqp_optimize(model, params, res);
    print_binding("binding_mps_max_iter", st, res);
    /* the same model and parameter object solved to optimality */
    hprqp_params_set_max_iter(params, 1000000);
    hprqp_params_set_stoptol(params, 1e-8);
    hprqp_params_set_check_iter(params, 100);
    hprqp_params_set_print_frequency(params, -1);
    hprqp_params_set_use_scaling(params, 1);
    hprqp_params_set_time_limit(params, 60.0);
    hprqp_params_set_spmv_mode_q(params, HPRQP_SPMV_MODE_AUTO);
    hprqp_params_set_spmv_mode_a(params, HPRQP_SPMV_MODE_AUTO);
    st = hprqp_optimize(model, params, res);
    print_binding("binding_mps_optimal", st, res);
    hprqp_results_destroy(res);
    hprqp_params_destroy(params);
    hprqp_model_destroy(model);
    return 0;
}

/* python/tests/test_hprqp.py:45-60 through the model API of hprqp.h:
 * min x0^2 + 2 x1^2 + 2 x0 + x1, x0 + x1 = 2, 0 <= x <= 10; the optimum is
 * x = (7/6, 5/6), objective 213/36. */
static int model_api_equality(void)
{
    size_t q_ptr[3] = {0, 1, 2}, q_ind[2] = {0, 1}, a_ptr[2] = {0, 2}, a_ind[2] = {0, 1};
    double q_val[2] = {2.0, 4.0}, a_val[2] = {1.0, 1.0};
    const double c[2] = {2.0, 1.0}, AL[1] = {2.0}, AU[1] = {2.0}, l[2] = {0.0, 0.0}, u[2] = {10.0, 10.0};
    HPRQPMatrixCSR Q = {2, 2, 2, q_ptr, q_ind, q_val, 0};
    HPRQPMatrixCSR A = {1, 2, 2, a_ptr, a_ind, a_val, 0};
    HPRQPModel model;
    HPRQPParams params;
    HPRQPResults res;
    model_init(&model);
    results_init(&res);
    HPRQPStatus st = build_from_QAbc(&Q, &A, c, AL, AU, l, u, 0.0, &model);
    if (st != HPRQP_STATUS_OK) return 2;
    params_set_defaults(&params);
    params.verbose = 0;
    params.use_gpu = 1;
    params.device_number = 0;
    params.stoptol = 1e-8;
    printf("gpu_backend_available value=%d\n", gpu_backend_available());
    st = solve(&model, &params, &res);
    printf("model_api_equality rc=%s status=%s iter=%zu objective=%.17g residual=%.17g gap=%.17g x0=%.17g x1=%.17g "
           "y0=%.17g\n",
           hprqp_status_name(st), hprqp_status_name((HPRQPStatus)res.status), res.iter, res.primal_obj, res.residuals,
           res.gap, res.x ? res.x[0] : NAN, res.x ? res.x[1] : NAN, res.y ? res.y[0] : NAN);
    /* the equality-row metric asked for, as solve_qp --gram request does
     * (examples/solve_qp.c:223-227): taken for the one equality row, the
     * same optimum; AUTO declines a plain QP */
    for (int mode = HPRQP_GRAM_AUTO; mode <= HPRQP_GRAM_REQUEST; ++mode) {
        HPRQPResults g;
        results_init(&g);
        params.gram_mode = (HPRQPGramMode)mode;
        st = solve(&model, &params, &g);
        printf("model_api_gram_%s rc=%s status=%s gram_active=%d objective=%.17g x0=%.17g x1=%.17g\n",
               mode == HPRQP_GRAM_AUTO ? "auto" : "request", hprqp_status_name(st),
               hprqp_status_name((HPRQPStatus)g.status), g.gram_active, g.primal_obj, g.x ? g.x[0] : NAN,
               g.x ? g.x[1] : NAN);
        results_free(&g);
    }
    params.gram_mode = HPRQP_GRAM_AUTO;
    /* the CPU path is the reference's, not this library's */
    params.use_gpu = 0;
    HPRQPResults cpu;
    results_init(&cpu);
    st = solve(&model, &params, &cpu);
    printf("model_api_cpu rc=%s\n", hprqp_status_name(st));
    results_free(&cpu);
    results_free(&res);
    model_free(&model);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s simple_qp.mps\n", argv[0]);
        return 64;
    }
    int rc = binding_unconstrained();
    if (!rc) rc = binding_mps_max_iter(argv[1]);
    if (!rc) rc = model_api_equality();
    printf("done rc=%d\n", rc);
    return rc;
}
