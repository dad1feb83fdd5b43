// Host control loop of the HPR method, shared by every backend of this
// library. It is the operator seam the reference calls HPRQPSolverBackendOps
// (src/shared/solver_backend.h:18-36): the loop owns the restart state machine
// and the termination test; the backend owns all vector state and runs the
// iteration, residuals, M-norm, sigma update and restart on the device.
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#include "shared/types.h"

namespace hprqp {

namespace polish {
class Owner;
}

class Backend {
public:
    virtual ~Backend() = default;
    virtual size_t n() const = 0;
    virtual size_t m() const = 0;
    virtual double sigma() const = 0;
    virtual void set_to_check(int to_check) = 0;
    // solver_backend.h:20-34
    virtual HPRQPStatus compute_residuals(HPRQPResiduals *res, size_t iter) = 0;
    virtual HPRQPStatus copy_output_iterates(double *x, double *w, double *z, double *y) = 0;
    virtual HPRQPStatus update_sigma(HPRQPRestart *restart, const HPRQPResiduals *res) = 0;
    virtual HPRQPStatus restart_apply(HPRQPRestart *restart) = 0;
    virtual double compute_m_norm() = 0;
    virtual HPRQPStatus iteration_step(size_t inner) = 0;
    // 1 when the equality-row (Gram) metric runs (results.gram_active)
    virtual int gram_active() const { return 0; }
};

// Restart bookkeeping, src/cpu/control.c:129-142, 231-305.
void restart_init(HPRQPRestart *r);
void restart_check(HPRQPRestart *r, size_t iter, size_t check_iter, double sigma);
void restart_post_iteration(HPRQPRestart *r, size_t iter, size_t check_iter, double gap);
// Scalar golden-section search of control.c:45-77 (non-diagonal Q).
double golden_search_sigma(double a, double b, double c, double d);
// The sigma blend + kappa correction of control.c:385-404, given the
// search estimate; returns the new sigma.
double sigma_from_estimate(double estimate, double sigma_old, const HPRQPRestart *r, const HPRQPResiduals *res);

// The loop of src/shared/solver_backend.c:338-569, runnable one body at a time.
// With `model` given and params->polishing set, check points also feed the
// host polishing worker (solver_backend.c:429, 467-479); the model must then
// outlive the loop.
class SolveLoop {
public:
    SolveLoop(Backend *backend, const HPRQPParams *params, const char *label, const HPRQPModel *model = nullptr);
    ~SolveLoop();
    HPRQPStatus begin();                    // iteration-0 residuals, header
    HPRQPStatus body();                     // one loop body
    bool finished() const { return finished_; }
    HPRQPStatus final_status() const { return final_status_; }
    size_t iter() const { return iter_; }
    const HPRQPResiduals &residuals() const { return res_; }
    const HPRQPRestart &restart() const { return restart_; }
    // Fill results (allocates x/y/z/w) as collect_results does (:289-336).
    HPRQPStatus collect(HPRQPResults *results);
    double elapsed() const;

private:
    Backend *b_;
    const HPRQPParams *p_;
    const char *label_;
    HPRQPResiduals res_;
    HPRQPRestart restart_;
    double start_ = 0.0;
    size_t iter_ = 0;
    bool started_ = false;
    bool finished_ = false;
    HPRQPStatus final_status_ = HPRQP_STATUS_CONTINUE;
    int first4_ = 1, first6_ = 1;
    size_t iter4_ = 0, iter6_ = 0;
    double time4_ = 0.0, time6_ = 0.0;
    const HPRQPModel *model_;
    std::unique_ptr<polish::Owner> polisher_;
    size_t polish_attempts_ = 0;
    bool polish_accepted_ = false;
    std::vector<double> polished_x_, polished_w_, polished_z_, polished_y_;  // an accepted candidate
    HPRQPStatus termination(size_t iter, double t) const;
    void end_polishing();
    void milestones(size_t iter, double t);
    void log_line(size_t iter, double t) const;
};

double wall_seconds();

}  // namespace hprqp
