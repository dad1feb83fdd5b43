// Host control loop and restart / sigma scalar logic. Semantics follow
// src/shared/solver_backend.c:119-569 and src/cpu/control.c:20-411.
#include "solve_loop.h"

#include "hprqp.h"
#include "polish_owner.h"

#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace hprqp {

double wall_seconds()
{
    struct timeval tv;
    if (gettimeofday(&tv, nullptr) != 0) return 0.0;
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

void restart_init(HPRQPRestart *r)
{
    std::memset(r, 0, sizeof(*r));
    r->first_restart = 1;
    r->save_gap = INFINITY;
    r->current_gap = INFINITY;
    r->last_gap = INFINITY;
    r->best_gap = INFINITY;
    r->best_sigma = 1.0;
    r->weighted_norm = INFINITY;
}

// control.c:231-282
void restart_check(HPRQPRestart *r, size_t iter, size_t check_iter, double sigma)
{
    if (check_iter == 0) return;
    r->restart_flag = 0;
    if (r->first_restart) {
        if (iter == check_iter) {
            r->first_restart = 0;
            r->restart_flag = 1;
            r->weighted_norm = r->current_gap;
            r->best_gap = r->current_gap;
            r->best_sigma = sigma;
        }
        return;
    }
    if (iter % check_iter != 0) return;
    if (r->current_gap < 0.0) r->current_gap = 1e-6;
    if (r->current_gap <= 0.2 * r->last_gap) {
        r->sufficient += 1;
        r->restart_flag = 1;
    }
    if (r->current_gap <= 0.8 * r->last_gap && r->current_gap > r->save_gap) {
        r->necessary += 1;
        r->restart_flag = 2;
    }
    double fact = r->weighted_norm > 0.0 && r->current_gap / r->weighted_norm > 1e-1 ? 0.5 : 0.2;
    if ((double)r->inner >= fact * (double)iter) {
        r->long_restart += 1;
        r->restart_flag = 3;
    }
    if (r->best_gap > r->current_gap) {
        r->best_gap = r->current_gap;
        r->best_sigma = sigma;
    }
    r->save_gap = r->current_gap;
}

// control.c:284-305
void restart_post_iteration(HPRQPRestart *r, size_t iter, size_t check_iter, double gap)
{
    bool block_end = check_iter > 0 && (iter + 1) % check_iter == 0;
    if (r->restart_flag > 0 || block_end) {
        if (r->restart_flag > 0) r->last_gap = gap;
        if (block_end) r->current_gap = gap;
    }
}

static double sigma_objective(double x, double a, double b, double c, double d)
{
    return a * x + b / x + c * x * x / (1.0 + d * x);
}

// control.c:45-77
double golden_search_sigma(double a, double b, double c, double d)
{
    double lo = 1e-12, hi = 1e12;
    const double g = (std::sqrt(5.0) - 1.0) / 2.0;
    const double g2 = (3.0 - std::sqrt(5.0)) / 2.0;
    double x1 = lo + g2 * (hi - lo), x2 = lo + g * (hi - lo);
    double f1 = sigma_objective(x1, a, b, c, d), f2 = sigma_objective(x2, a, b, c, d);
    for (int it = 0; it < 256; ++it) {
        if (std::fabs(hi - lo) <= 1e-13 * std::fmax(1.0, std::fmax(std::fabs(lo), std::fabs(hi)))) break;
        if (f1 > f2) {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + g * (hi - lo);
            f2 = sigma_objective(x2, a, b, c, d);
        } else {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = lo + g2 * (hi - lo);
            f1 = sigma_objective(x1, a, b, c, d);
        }
    }
    return 0.5 * (lo + hi);
}

// control.c:385-404
double sigma_from_estimate(double est, double sigma_old, const HPRQPRestart *r, const HPRQPResiduals *res)
{
    est = clampd(est, 1e-12, 1e12);
    double fact = (r->best_gap > 0.0 && std::isfinite(r->best_gap)) ? std::exp(-0.05 * (r->current_gap / r->best_gap))
                                                                    : 1.0;
    double cand = std::exp(fact * std::log(est) +
                           (1.0 - fact) * std::log(r->best_sigma > 0.0 ? r->best_sigma : sigma_old));
    double t1 = std::fmax(std::fmin(res->err_Rd_org_bar, res->err_Rp_org_bar),
                          std::fmin(res->rel_gap_bar, r->current_gap));
    double kappa = 1.0;
    if (t1 <= 9e-10 && res->err_Rp_org_bar > 0.0) {
        double ratio = res->err_Rd_org_bar / res->err_Rp_org_bar;
        kappa = t1 > 5e-10 ? clampd(std::sqrt(ratio), 1e-2, 100.0) : clampd(ratio, 1e-2, 100.0);
    }
    return clampd(kappa * cand, 1e-12, 1e12);
}

// solver_backend.c:119-132
static size_t adaptive_print_step(size_t iter)
{
    size_t step = 10;
    if (iter == 0) return step;
    while (iter / step >= 100 && step <= SIZE_MAX / 10) step *= 10;
    return step;
}

static bool print_current(const HPRQPParams *p, size_t iter, double t)
{
    if (p->print_frequency == -1 && (iter % adaptive_print_step(iter) == 0 || iter == p->max_iter || t > p->time_limit))
        return true;
    if (p->print_frequency > 0 &&
        (iter % (size_t)p->print_frequency == 0 || iter == p->max_iter || t > p->time_limit))
        return true;
    return false;
}

static int check_next(const HPRQPParams *p, const HPRQPRestart *r, size_t next)
{
    int chk = 0;
    if (p->check_iter > 0 && next % p->check_iter == 0) chk = 1;
    if (r->restart_flag > 0) chk = 1;
    if (p->print_frequency == -1 && next % adaptive_print_step(next) == 0) chk = 1;
    if (p->print_frequency > 0 && next % (size_t)p->print_frequency == 0) chk = 1;
    return chk;
}

SolveLoop::SolveLoop(Backend *backend, const HPRQPParams *params, const char *label, const HPRQPModel *model)
    : b_(backend), p_(params), label_(label), model_(model)
{
    std::memset(&res_, 0, sizeof(res_));
    restart_init(&restart_);
}

SolveLoop::~SolveLoop() = default;

// solver_backend.c:193-211 (check_termination_status)
HPRQPStatus SolveLoop::termination(size_t iter, double t) const
{
    if (res_.KKTx_and_gap_org_bar < p_->stoptol) return HPRQP_STATUS_OPTIMAL;
    if (iter == p_->max_iter) return HPRQP_STATUS_MAX_ITER;
    if (t > p_->time_limit) return HPRQP_STATUS_TIME_LIMIT;
    return HPRQP_STATUS_CONTINUE;
}

// solver_backend.c:547, 567 (hprqp_polish_destroy at the end of the solve):
// a worker still running is stopped and its files removed; the counts stay
// for the results.
void SolveLoop::end_polishing()
{
    if (!polisher_) return;
    polish_attempts_ = polisher_->attempts();
    polish_accepted_ = polisher_->accepted();
    polisher_.reset();
}

double SolveLoop::elapsed() const
{
    double now = wall_seconds();
    if (now <= 0.0 || start_ <= 0.0) return 0.0;
    return now - start_;
}

void SolveLoop::milestones(size_t iter, double t)
{
    if (first4_ && res_.KKTx_and_gap_org_bar < 1e-4) {
        iter4_ = iter;
        time4_ = t;
        first4_ = 0;
    }
    if (first6_ && res_.KKTx_and_gap_org_bar < 1e-6) {
        iter6_ = iter;
        time6_ = t;
        first6_ = 0;
    }
}

void SolveLoop::log_line(size_t iter, double t) const
{
    std::printf("%5zu    %3.2e    %3.2e    %+7.6e    %+7.6e    %3.2e    %3.2e    %6.2f\n", iter, res_.err_Rp_org_bar,
                res_.err_Rd_org_bar, res_.primal_obj_bar, res_.dual_obj_bar, res_.rel_gap_bar, b_->sigma(), t);
}

// solver_backend.c:375-431
HPRQPStatus SolveLoop::begin()
{
    start_ = wall_seconds();
    HPRQPStatus st = b_->compute_residuals(&res_, 0);
    if (st != HPRQP_STATUS_OK) return st;
    milestones(0, elapsed());
    if (p_->verbose) {
        std::printf("HPRQP SOLVER starts (%s)...\n", label_ ? label_ : "backend mode");
        std::printf(" iter     errRp        errRd         p_obj           d_obj          gap        sigma       time\n");
        log_line(0, elapsed());
    }
    // solver_backend.c:429
    if (model_) polisher_.reset(polish::Owner::create(model_, p_));
    restart_.best_sigma = b_->sigma();
    iter_ = 0;
    started_ = true;
    return HPRQP_STATUS_OK;
}

// solver_backend.c:432-513 (one pass of the for-body)
HPRQPStatus SolveLoop::body()
{
    if (!started_) {
        HPRQPStatus st = begin();
        if (st != HPRQP_STATUS_OK) return st;
    }
    if (finished_) return HPRQP_STATUS_OK;
    const size_t iter = iter_;
    bool print_yes = false;
    double t = elapsed();
    HPRQPStatus st;
    if (iter > 0) {
        print_yes = print_current(p_, iter, t);
        if ((p_->check_iter > 0 && iter % p_->check_iter == 0) || print_yes) {
            st = b_->compute_residuals(&res_, iter);
            if (st != HPRQP_STATUS_OK) return st;
            milestones(iter, elapsed());
        } else {
            res_.is_updated = 0;
        }
    }
    final_status_ = termination(iter, t);
    // solver_backend.c:467-479: a certified candidate from the polishing
    // worker ends the solve at this check point
    if (polisher_ && final_status_ == HPRQP_STATUS_CONTINUE && res_.is_updated) {
        const size_t n = b_->n(), m = b_->m();
        if (polished_x_.size() != n || polished_y_.size() != m) {  // written only when a candidate is accepted
            polished_x_.resize(n);
            polished_w_.resize(n);
            polished_z_.resize(n);
            polished_y_.resize(m);
        }
        HPRQPResiduals polished;
        if (polisher_->poll(b_, res_, iter, t, polished_x_.data(), polished_w_.data(), polished_z_.data(),
                            polished_y_.data(), &polished)) {
            res_ = polished;
            final_status_ = HPRQP_STATUS_OPTIMAL;
            finished_ = true;
            end_polishing();
            return HPRQP_STATUS_OK;
        }
        t = elapsed();
        final_status_ = termination(iter, t);
    }

    restart_check(&restart_, iter, p_->check_iter, b_->sigma());
    st = b_->update_sigma(&restart_, &res_);
    if (st != HPRQP_STATUS_OK) return st;
    st = b_->restart_apply(&restart_);
    if (st != HPRQP_STATUS_OK) return st;
    if (p_->verbose && (print_yes || final_status_ != HPRQP_STATUS_CONTINUE)) log_line(iter, elapsed());
    if (final_status_ != HPRQP_STATUS_CONTINUE) {
        finished_ = true;
        end_polishing();
        return HPRQP_STATUS_OK;
    }
    b_->set_to_check(check_next(p_, &restart_, iter + 1));
    st = b_->iteration_step(restart_.inner);
    if (st != HPRQP_STATUS_OK) return st;
    if (restart_.restart_flag > 0 || (p_->check_iter > 0 && (iter + 1) % p_->check_iter == 0)) {
        double gap = b_->compute_m_norm();
        if (!std::isfinite(gap)) return HPRQP_STATUS_INVALID_ARGUMENT;
        restart_post_iteration(&restart_, iter, p_->check_iter, gap);
    }
    restart_.inner += 1;
    iter_ = iter + 1;
    return HPRQP_STATUS_OK;
}

// solver_backend.c:289-336 plus the output copy of :536-561
HPRQPStatus SolveLoop::collect(HPRQPResults *results)
{
    const size_t n = b_->n(), m = b_->m();
    const double t = elapsed();
    results_free(results);
    results->x = n ? static_cast<double *>(std::calloc(n, sizeof(double))) : nullptr;
    results->w = n ? static_cast<double *>(std::calloc(n, sizeof(double))) : nullptr;
    results->z = n ? static_cast<double *>(std::calloc(n, sizeof(double))) : nullptr;
    results->y = m ? static_cast<double *>(std::calloc(m, sizeof(double))) : nullptr;
    if ((n && (!results->x || !results->w || !results->z)) || (m && !results->y)) {
        results_free(results);
        return HPRQP_STATUS_ALLOCATION_FAILED;
    }
    if (polish_accepted_ && polished_x_.size() == n) {
        // solver_backend.c:472-474: the accepted candidate is the output
        std::copy(polished_x_.begin(), polished_x_.end(), results->x);
        std::copy(polished_w_.begin(), polished_w_.end(), results->w);
        std::copy(polished_z_.begin(), polished_z_.end(), results->z);
        std::copy(polished_y_.begin(), polished_y_.end(), results->y);
    } else {
        HPRQPStatus st = b_->copy_output_iterates(results->x, results->w, results->z, results->y);
        if (st != HPRQP_STATUS_OK) return st;
    }
    results->iter = iter_;
    results->time = t;
    results->iter_4 = iter4_ == 0 ? iter_ : iter4_;
    results->iter_6 = iter6_ == 0 ? iter_ : iter6_;
    results->time_4 = time4_ == 0.0 ? t : time4_;
    results->time_6 = time6_ == 0.0 ? t : time6_;
    results->primal_obj = res_.primal_obj_bar;
    results->residuals = res_.KKTx_and_gap_org_bar;
    results->gap = res_.rel_gap_bar;
    results->status = finished_ ? final_status_ : HPRQP_STATUS_CONTINUE;
    results->polish_attempts = polisher_ ? polisher_->attempts() : polish_attempts_;
    results->polish_accepted = polisher_ ? (int)polisher_->accepted() : (int)polish_accepted_;
    results->gram_active = b_->gram_active();
    return HPRQP_STATUS_OK;
}

}  // namespace hprqp
