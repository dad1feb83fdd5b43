// Equality-row (Gram) metric of the HPR iteration on one GPU.
//
// With the rows of A split as [E; C] (E the metric rows), G = E E' and
// beta >= ||C||_2^2 (majorised by the largest row sum times the largest
// column sum of |C|), the metric replaces the y update of the E rows by
//   y_bar_E = y_E - G^{-1} (E t - b_E) / (2 sigma)
// and runs the other rows with lambda_A = 2 beta
// (src/gpu/qp_row_metric.cu:55-63, src/gpu/row_metric.cu:46-55). Two forms:
//   * plain QPs (gram_mode REQUEST): E = the rows with AL == AU finite,
//     sigma from the metric norm of b (qp_row_metric.cu:94-149);
//   * row-cone models: E = the linear prefix, all homogeneous equalities,
//     diagonal Q, sigma / sqrt(beta) (row_metric.cu:68-130); AUTO takes this
//     form for the grid-stencil models of src/shared/gram_policy.cpp.
//
// The reference factors G with a sparse direct solver. Here G (at most
// kGramMaxRows rows) is factored on the device by a blocked FP64 Cholesky,
// the factor inverted to W = L^{-1}, and each step solves with two
// triangular matrix-vector products, u = W r and x = W' u: row-major W and
// W' stream from HBM with every workgroup busy, instead of a chain of
// dependent triangular sweeps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "shared/types.h"

namespace hprqp {

constexpr int kGramMaxRows = 16384;

// The scaled model as the device holds it, copied to the host for the setup.
struct GramInputs {
    int n = 0, m = 0;
    const int *rp = nullptr, *ci = nullptr;  // A, CSR
    const double *val = nullptr;
    const double *AL = nullptr, *AU = nullptr, *c = nullptr;
    const double *soc_rhs = nullptr;  // null without row cones
    size_t num_var_cones = 0, num_row_cones = 0;
    int lin_m = 0;                    // rows before the first row cone
    int Q_is_diag = 0, noC = 0;
};

class GramMetric {
public:
    GramMetric() = default;
    GramMetric(const GramMetric &) = delete;
    GramMetric &operator=(const GramMetric &) = delete;
    ~GramMetric();

    // Builds the metric for `in` on `stream`. Returns OK with kind() == 0 when
    // the model is declined (reason set); device errors return their status.
    // On acceptance *sigma and *lambda_A hold the metric's values.
    HPRQPStatus setup(const GramInputs &in, hipStream_t stream, double *sigma, double *lambda_A,
                      const char **reason);
    int kind() const { return kind_; }  // 0 off, 1 plain-QP rows, 2 row-cone prefix
    int rows() const { return e_; }
    double beta() const { return beta_; }
    const int *map() const { return d_map_; }  // row -> metric index or -1 (m entries)
    double *rhs() { return d_r_; }             // E t - b_E, written by the y step
    const double *b() const { return d_b_; }

    // Metric y_bar on the E rows from rhs(): step = G^{-1} rhs, then
    // y_bar = y - step / (2 sigma), dy, and the Halpern y (weights a, b).
    HPRQPStatus finish_y(hipStream_t s, double *y, double *y_bar, double *dy, const double *last_y, double sigma,
                         double a, double b);
    // 2 ||E' dy||^2 + 2 beta ||dy_rest||^2 (read back to the host).
    HPRQPStatus quadratic(hipStream_t s, const double *dy, double *out);
    // Device vectors by name for inspection: "gram_r", "gram_step".
    size_t read(const char *name, double *out, size_t cap, hipStream_t s) const;

private:
    int kind_ = 0, e_ = 0, n_ = 0, m_ = 0, ld_ = 0;
    double beta_ = 0.0;
    double *d_W_ = nullptr, *d_WT_ = nullptr;  // W = L^{-1} and W', row-major, ld_ x ld_
    double *d_r_ = nullptr, *d_u_ = nullptr, *d_step_ = nullptr, *d_b_ = nullptr;
    int *d_rows_ = nullptr, *d_map_ = nullptr;
    int *d_et_rp_ = nullptr, *d_et_ci_ = nullptr;  // E' (n rows, column = row of A)
    double *d_et_val_ = nullptr;
    double *d_part_ = nullptr;  // per-workgroup partial sums of the quadratic
    int part_blocks_ = 0;
    std::vector<double> part_h_;
    HPRQPStatus factor(const std::vector<double> &G, hipStream_t s, bool *spd);
    HPRQPStatus solve(const std::vector<double> &rhs, std::vector<double> &x, hipStream_t s);
    void release();
};

// src/shared/gram_policy.cpp:6-50: the AUTO rule on the built, unscaled model.
bool gram_auto_eligible(const HPRQPModel *model);

}  // namespace hprqp
