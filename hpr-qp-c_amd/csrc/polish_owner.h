// Owner side of the reference's host polishing hook (src/shared/polish.h,
// src/shared/polish.cpp, src/shared/polish/*.hpp): while the GPU iterates,
// an eligible check point is exported to a separate CPU worker process
// (the reference's native `hprqp_polish_worker`, or a Python worker when
// `polish_python` is set); a worker that publishes a certified candidate
// ends the solve as OPTIMAL with that candidate. The worker itself is the
// reference's and runs unchanged; this file implements the owner's rule,
// process handling and file protocol. Host code only: no GPU memory is
// touched, the snapshot comes through Backend::copy_output_iterates.
#pragma once

#include <array>
#include <chrono>
#include <cstddef>
#include <cstdint>
#include <string>
#include <sys/types.h>
#include <vector>

#include "shared/types.h"

namespace hprqp {
class Backend;
}

namespace hprqp::polish {

// Launch rule of src/shared/polish/progress.hpp:11-37: a check point is
// eligible from iteration 7 * check_iter on, when the KKT error is above
// the tolerance and has crossed the next unconsumed decade of either the
// total target (floor 10 tol) or the equation target (max(Rp, Rd), floor
// min(10 tol, 64 eps)). A launch consumes every decade the point has
// crossed; busy periods coalesce crossings.
struct Milestones {
    double total_target = 1e-2;
    double equation_target = 1e-2;
    unsigned launches = 0;
    static double equation_floor(double tol);
    bool eligible(size_t iter, size_t check_iter, double rp, double rd, double kkt, double tol) const;
    void launched(double rp, double rd, double kkt, double tol);
};

// SHA-256 (FIPS 180-4) of a byte string and of a file's contents.
std::array<unsigned char, 32> sha256(const void *data, size_t len);
std::array<unsigned char, 32> sha256_file(const std::string &path);

// One worker process in a process group of its own, with a wall-clock
// budget (src/shared/polish/async_process.hpp:17-95): stdout and stderr go
// to a new log file, BLAS/OpenMP thread counts are pinned to 1 in the
// child's environment only, and a timed-out or cancelled worker's whole
// group is killed and reaped.
class Worker {
public:
    enum State { IDLE, RUNNING, EXITED, TIMED_OUT, CANCELLED };
    Worker() = default;
    Worker(const Worker &) = delete;
    Worker &operator=(const Worker &) = delete;
    ~Worker() { cancel(); }
    void start(const std::vector<std::string> &argv, const std::string &log, double seconds);
    State poll();
    void cancel();
    bool running() const { return pid_ > 0; }
    double age() const;
    State state() const { return state_; }
    int exit_code() const { return exit_code_; }

private:
    pid_t pid_ = -1;
    State state_ = IDLE;
    int exit_code_ = -1;
    double budget_ = 0.0;
    std::chrono::steady_clock::time_point started_{};
    void stop(State why);
};

// The worker's published result (src/shared/polish/cpu_publication.hpp).
struct Certificate {
    double tolerance, primal_feasibility, dual_feasibility, gap, primal_objective, dual_objective;
};
// READY holds "HPRCPU01\n<token>\n" and nothing else.
bool ready_matches(const std::string &path, const std::string &token);
// certificate.bin: "HPRCPU01", 16-byte id (the token), n, m, SHA-256 of
// candidate.bin, then tolerance, Rp, Rd, gap, primal and dual objective;
// throws unless every field is consistent with the request.
Certificate read_certificate(const std::string &episode, const std::string &token, uint64_t n, uint64_t m,
                             double tolerance);
// candidate.bin: "HPRPOL01", n, m, then x, w, z (n each), y, s (m each).
struct Candidate {
    std::vector<double> x, w, z, y, s;
};
Candidate read_candidate(const std::string &path, size_t n, size_t m);
// model.bin ("HPRNAT01") and snapshot.bin ("HPRSOLS1"), written read-only.
void write_model(const HPRQPModel &model, const std::string &path);
void write_snapshot(const std::string &path, size_t n, size_t m, const double *x, const double *y, const double *z,
                    const double *w);

// Worker lookup (src/shared/polish/worker_path.hpp:10-40): the override
// path if given; otherwise `hprqp_polish_worker` (or `polishing/worker.py`
// for a Python worker) beside each anchor (this library, the executable),
// then `libexec/hprqp/hprqp_polish_worker` (`share/hprqp/polishing/
// worker.py`) up to three directories above it. A native worker must be
// executable. Returns an empty string when none is found.
std::string find_worker(const char *override_path, const std::vector<std::string> &anchors, bool python);
std::vector<std::string> default_anchors();

class Owner {
public:
    // nullptr when polishing is off, the model has variable cones, or the
    // worker cannot be set up (as hprqp_polish_create, polish.cpp:140-143).
    static Owner *create(const HPRQPModel *model, const HPRQPParams *params);
    ~Owner();
    // One check point (polish.cpp:80-133). Returns true when a certified
    // candidate was accepted: x, w, z (n) and y (m) then hold it and *out its
    // residuals. Errors disable polishing for the rest of the solve.
    bool poll(Backend *backend, const HPRQPResiduals &base, size_t iter, double seconds, double *x, double *w,
              double *z, double *y, HPRQPResiduals *out);
    size_t attempts() const { return attempts_; }
    bool accepted() const { return accepted_; }

private:
    Owner(const HPRQPModel *model, const HPRQPParams *params);
    bool poll_impl(Backend *backend, const HPRQPResiduals &base, size_t iter, double seconds, double *x, double *w,
                   double *z, double *y, HPRQPResiduals *out);
    void event(const char *what, size_t iter);

    HPRQPParams params_;
    const HPRQPModel *model_;
    std::string worker_, python_, directory_;
    Milestones progress_;
    Worker child_;
    std::string run_, episode_, token_;
    bool pending_ = false, accepted_ = false, disabled_ = false;
    size_t attempts_ = 0;
    double child_seconds_ = 0.0;
};

}  // namespace hprqp::polish
