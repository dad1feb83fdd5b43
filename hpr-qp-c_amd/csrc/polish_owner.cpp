// Owner side of the host polishing hook; the protocol and rules are those
// of src/shared/polish.cpp and src/shared/polish/*.hpp (cited per function).
#include "polish_owner.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <fcntl.h>
#include <filesystem>
#include <fstream>
#include <limits>
#include <signal.h>
#include <spawn.h>
#include <sstream>
#include <stdexcept>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include "solve_loop.h"

extern char **environ;

namespace hprqp::polish {

namespace fs = std::filesystem;
using Clock = std::chrono::steady_clock;

// ------------------------------------------------------------ milestones
// progress.hpp:16-18
double Milestones::equation_floor(double tol) { return std::min(10.0 * tol, 64.0 * std::numeric_limits<double>::epsilon()); }

// progress.hpp:19-26
bool Milestones::eligible(size_t iter, size_t check_iter, double rp, double rd, double kkt, double tol) const
{
    if (iter < 7 * std::max<size_t>(1, check_iter) || !(tol > 0.0)) return false;
    if (!std::isfinite(rp) || !std::isfinite(rd) || !std::isfinite(kkt)) return false;
    if (rp < 0.0 || rd < 0.0 || kkt <= tol) return false;
    const bool total = total_target >= 10.0 * tol && kkt <= total_target;
    const bool equations = equation_target >= equation_floor(tol) && std::max(rp, rd) <= equation_target;
    return total || equations;
}

namespace {
// progress.hpp:27-32: every decade the value has crossed is used up; a
// target that falls below its floor by rounding lands on the floor, one
// that passes the floor for good becomes 0 (never eligible again).
void consume(double &target, double value, double floor)
{
    while (target >= floor && value <= target) target *= 0.1;
    if (target < floor * (1.0 - 1e-12)) target = 0.0;
    else if (target < floor) target = floor;
}
}  // namespace

// progress.hpp:33-36
void Milestones::launched(double rp, double rd, double kkt, double tol)
{
    ++launches;
    consume(total_target, kkt, 10.0 * tol);
    consume(equation_target, std::max(rp, rd), equation_floor(tol));
}

// ----------------------------------------------------------------- SHA-256
namespace {
constexpr uint32_t kSha[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

inline uint32_t rotr(uint32_t v, int s) { return (v >> s) | (v << (32 - s)); }

class Sha256 {
public:
    void update(const unsigned char *p, size_t len)
    {
        bytes_ += len;
        while (len) {
            const size_t take = std::min(len, sizeof(buf_) - fill_);
            std::memcpy(buf_ + fill_, p, take);
            fill_ += take;
            p += take;
            len -= take;
            if (fill_ == sizeof(buf_)) {
                block(buf_);
                fill_ = 0;
            }
        }
    }
    std::array<unsigned char, 32> digest()
    {
        const uint64_t bits = bytes_ * 8;
        const unsigned char one = 0x80, zero = 0;
        update(&one, 1);
        while (fill_ != 56) update(&zero, 1);
        unsigned char len[8];
        for (int i = 0; i < 8; ++i) len[i] = (unsigned char)(bits >> (56 - 8 * i));
        update(len, 8);
        std::array<unsigned char, 32> out{};
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 4; ++j) out[4 * i + j] = (unsigned char)(h_[i] >> (24 - 8 * j));
        return out;
    }

private:
    uint32_t h_[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    unsigned char buf_[64];
    size_t fill_ = 0;
    uint64_t bytes_ = 0;
    void block(const unsigned char *p)
    {
        uint32_t w[64];
        for (int t = 0; t < 16; ++t)
            w[t] = (uint32_t)p[4 * t] << 24 | (uint32_t)p[4 * t + 1] << 16 | (uint32_t)p[4 * t + 2] << 8 | p[4 * t + 3];
        for (int t = 16; t < 64; ++t) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], h = h_[7];
        for (int t = 0; t < 64; ++t) {
            const uint32_t t1 = h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + kSha[t] + w[t];
            const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            h = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h_[0] += a;
        h_[1] += b;
        h_[2] += c;
        h_[3] += d;
        h_[4] += e;
        h_[5] += f;
        h_[6] += g;
        h_[7] += h;
    }
};
}  // namespace

std::array<unsigned char, 32> sha256(const void *data, size_t len)
{
    Sha256 s;
    s.update(static_cast<const unsigned char *>(data), len);
    return s.digest();
}

std::array<unsigned char, 32> sha256_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("candidate missing");
    Sha256 s;
    std::vector<char> buf(1 << 16);
    while (f) {
        f.read(buf.data(), (std::streamsize)buf.size());
        if (f.gcount() > 0) s.update(reinterpret_cast<const unsigned char *>(buf.data()), (size_t)f.gcount());
    }
    if (!f.eof()) throw std::runtime_error("candidate read failure");
    return s.digest();
}

// ------------------------------------------------------------------ worker
// async_process.hpp:30-71
void Worker::start(const std::vector<std::string> &argv, const std::string &log, double seconds)
{
    if (running() || argv.empty() || !(seconds > 0.0)) throw std::runtime_error("invalid worker start");
    static const char *const kThreadVars[] = {"OPENBLAS_NUM_THREADS=", "OMP_NUM_THREADS=", "MKL_NUM_THREADS=",
                                              "BLIS_NUM_THREADS=", "VECLIB_MAXIMUM_THREADS="};
    std::vector<std::string> env;
    for (char **e = environ; e && *e; ++e) {
        bool pinned = false;
        for (const char *v : kThreadVars)
            if (std::strncmp(*e, v, std::strlen(v)) == 0) pinned = true;
        if (!pinned) env.emplace_back(*e);
    }
    for (const char *v : kThreadVars) env.push_back(std::string(v) + "1");
    std::vector<char *> envp, args;
    for (auto &s : env) envp.push_back(s.data());
    envp.push_back(nullptr);
    std::vector<std::string> argcopy(argv);
    for (auto &s : argcopy) args.push_back(s.data());
    args.push_back(nullptr);

    posix_spawn_file_actions_t fa;
    posix_spawnattr_t at;
    if (posix_spawn_file_actions_init(&fa)) throw std::runtime_error("spawn actions");
    if (posix_spawnattr_init(&at)) {
        posix_spawn_file_actions_destroy(&fa);
        throw std::runtime_error("spawn attributes");
    }
    int rc = posix_spawn_file_actions_addopen(&fa, STDOUT_FILENO, log.c_str(), O_WRONLY | O_CREAT | O_EXCL, 0600);
    if (!rc) rc = posix_spawn_file_actions_adddup2(&fa, STDOUT_FILENO, STDERR_FILENO);
    if (!rc) rc = posix_spawnattr_setflags(&at, POSIX_SPAWN_SETPGROUP);
    if (!rc) rc = posix_spawnattr_setpgroup(&at, 0);  // the worker leads a group of its own
    started_ = Clock::now();
    budget_ = seconds;
    pid_t pid = -1;
    if (!rc) rc = posix_spawnp(&pid, args[0], &fa, &at, args.data(), envp.data());
    posix_spawn_file_actions_destroy(&fa);
    posix_spawnattr_destroy(&at);
    if (rc) {
        pid_ = -1;
        state_ = IDLE;
        throw std::runtime_error("posix_spawn error " + std::to_string(rc));
    }
    pid_ = pid;
    state_ = RUNNING;
    exit_code_ = -1;
}

double Worker::age() const
{
    return state_ == IDLE ? 0.0 : std::chrono::duration<double>(Clock::now() - started_).count();
}

// async_process.hpp:72-86: a worker past its budget is stopped even if
// its owner polls late.
Worker::State Worker::poll()
{
    if (!running()) return state_;
    int status = 0;
    const pid_t r = waitpid(pid_, &status, WNOHANG);
    if (r == pid_) {
        pid_ = -1;
        state_ = EXITED;
        exit_code_ = WIFEXITED(status) ? WEXITSTATUS(status) : 128 + WTERMSIG(status);
    } else if (r < 0 && errno != EINTR) {
        if (errno != ECHILD) throw std::runtime_error("worker waitpid");
        pid_ = -1;
        state_ = EXITED;
        exit_code_ = -1;
    }
    if (r == 0 && age() >= budget_) stop(TIMED_OUT);
    return state_;
}

void Worker::cancel()
{
    if (running()) stop(CANCELLED);
}

// async_process.hpp:88-94: SIGKILL to the worker's own group, then reap.
void Worker::stop(State why)
{
    if (pid_ <= 0) return;
    kill(-pid_, SIGKILL);
    int status = 0;
    while (waitpid(pid_, &status, 0) < 0 && errno == EINTR) {
    }
    pid_ = -1;
    state_ = why;
    exit_code_ = 137;
}

// ------------------------------------------------------------------- files
namespace {
template <class T>
void put(std::ofstream &f, const T *v, size_t count)
{
    if (count) f.write(reinterpret_cast<const char *>(v), (std::streamsize)(sizeof(T) * count));
}

std::string hex(const unsigned char *p, size_t len)
{
    static const char *digits = "0123456789abcdef";
    std::string s;
    for (size_t i = 0; i < len; ++i) {
        s += digits[p[i] >> 4];
        s += digits[p[i] & 15];
    }
    return s;
}

// polish.cpp:19-23: a JSON string; control characters are refused
std::string json_string(const std::string &s)
{
    std::string out = "\"";
    for (unsigned char c : s) {
        if (c < 32) throw std::runtime_error("control character in path");
        if (c == '\\' || c == '"') out += '\\';
        out += (char)c;
    }
    return out + '"';
}

// polish.cpp:24-29: 16 bytes from the kernel's entropy pool, as hex
std::string new_token()
{
    std::ifstream f("/dev/urandom", std::ios::binary);
    unsigned char b[16];
    if (!f.read(reinterpret_cast<char *>(b), sizeof(b))) throw std::runtime_error("entropy unavailable");
    return hex(b, sizeof(b));
}
}  // namespace

// async_process.hpp:97-101
bool ready_matches(const std::string &path, const std::string &token)
{
    std::ifstream f(path, std::ios::binary);
    std::string a, b;
    if (!std::getline(f, a) || !std::getline(f, b)) return false;
    return a == "HPRCPU01" && b == token && f.peek() == EOF;
}

// cpu_publication.hpp:42-66
Certificate read_certificate(const std::string &episode, const std::string &token, uint64_t n, uint64_t m,
                             double tolerance)
{
    std::ifstream f(episode + "/certificate.bin", std::ios::binary);
    char magic[8];
    unsigned char id[16];
    uint64_t dims[2];
    unsigned char digest[32];
    double v[6];
    if (!f.read(magic, 8) || !f.read(reinterpret_cast<char *>(id), 16) || !f.read(reinterpret_cast<char *>(dims), 16) ||
        !f.read(reinterpret_cast<char *>(digest), 32) || !f.read(reinterpret_cast<char *>(v), 48) || f.peek() != EOF)
        throw std::runtime_error("CPU certificate length");
    if (std::memcmp(magic, "HPRCPU01", 8) || hex(id, 16) != token || dims[0] != n || dims[1] != m)
        throw std::runtime_error("CPU certificate identity");
    for (double x : v)
        if (!std::isfinite(x)) throw std::runtime_error("nonfinite CPU certificate");
    if (!(tolerance > 0.0) || v[0] != tolerance) throw std::runtime_error("CPU certificate tolerance");
    for (int i = 1; i <= 3; ++i)
        if (v[i] < 0.0 || v[i] > tolerance) throw std::runtime_error("CPU certificate accuracy");
    const double gap = std::abs(v[4] - v[5]) / (1.0 + std::fmax(std::abs(v[4]), std::abs(v[5])));
    if (gap > tolerance || std::abs(gap - v[3]) > 1e-14) throw std::runtime_error("CPU certificate gap");
    if (n > (uint64_t)INT32_MAX || m > (uint64_t)INT32_MAX ||
        fs::file_size(episode + "/candidate.bin") != 24 + 8 * (3 * n + 2 * m))
        throw std::runtime_error("CPU candidate length");
    if (std::memcmp(sha256_file(episode + "/candidate.bin").data(), digest, sizeof(digest)))
        throw std::runtime_error("CPU certificate payload mismatch");
    return {v[0], v[1], v[2], v[3], v[4], v[5]};
}

// polish.cpp:45-57
Candidate read_candidate(const std::string &path, size_t n, size_t m)
{
    std::ifstream f(path, std::ios::binary);
    f.exceptions(std::ios::badbit | std::ios::failbit);
    char magic[8];
    uint64_t dims[2];
    f.read(magic, 8);
    f.read(reinterpret_cast<char *>(dims), 16);
    if (std::memcmp(magic, "HPRPOL01", 8) || dims[0] != n || dims[1] != m)
        throw std::runtime_error("candidate dimensions");
    Candidate c{std::vector<double>(n), std::vector<double>(n), std::vector<double>(n), std::vector<double>(m),
                std::vector<double>(m)};
    for (auto *v : {&c.x, &c.w, &c.z, &c.y, &c.s}) {
        if (!v->empty()) f.read(reinterpret_cast<char *>(v->data()), (std::streamsize)(8 * v->size()));
        for (double a : *v)
            if (!std::isfinite(a)) throw std::runtime_error("candidate not finite");
    }
    return c;
}

// polish.cpp:33-44: Q then A as (nrow, ncol, nnz, row_ptr, col_ind,
// values) with 64-bit sizes, then c, l, u, AL, AU, the cone right-hand
// sides (zeros without row cones), the objective constant and the row-cone
// offsets.
void write_model(const HPRQPModel &md, const std::string &path)
{
    static_assert(sizeof(size_t) == 8, "the polishing files use 64-bit sizes");
    std::ofstream f(path, std::ios::binary);
    f.exceptions(std::ios::badbit | std::ios::failbit);
    f.write("HPRNAT01", 8);
    for (const HPRQPMatrixCSR *M : {&md.Q, &md.A}) {
        const uint64_t d[3] = {M->nrow, M->ncol, M->nnz};
        put(f, d, 3);
        put(f, M->row_ptr, M->nrow + 1);
        put(f, M->col_ind, M->nnz);
        put(f, M->values, M->nnz);
    }
    const size_t n = md.Q.ncol, r = md.A.nrow;
    put(f, md.c, n);
    put(f, md.l, n);
    put(f, md.u, n);
    put(f, md.AL, r);
    put(f, md.AU, r);
    if (md.soc_rhs) {
        put(f, md.soc_rhs, r);
    } else {
        const std::vector<double> zero(r, 0.0);
        put(f, zero.data(), r);
    }
    put(f, &md.obj_constant, 1);
    const uint64_t cones = md.num_row_cones;
    put(f, &cones, 1);
    if (cones) put(f, md.row_cone_offsets, cones + 1);
    f.close();
    if (chmod(path.c_str(), 0400)) throw std::runtime_error("model file permissions");
}

// polish.cpp:113-116: dimensions (n, m, n, n), then x, y, z, w
void write_snapshot(const std::string &path, size_t n, size_t m, const double *x, const double *y, const double *z,
                    const double *w)
{
    std::ofstream f(path, std::ios::binary);
    f.exceptions(std::ios::badbit | std::ios::failbit);
    f.write("HPRSOLS1", 8);
    const uint64_t dims[4] = {n, m, n, n};
    put(f, dims, 4);
    put(f, x, n);
    put(f, y, m);
    put(f, z, n);
    put(f, w, n);
    f.close();
    if (chmod(path.c_str(), 0400)) throw std::runtime_error("snapshot permissions");
}

// ------------------------------------------------------------ worker lookup
// worker_path.hpp:10-40
std::string find_worker(const char *override_path, const std::vector<std::string> &anchors, bool python)
{
    std::vector<fs::path> candidates;
    if (override_path) {
        candidates.emplace_back(override_path);
    } else {
        for (const auto &a : anchors) {
            fs::path dir = fs::path(a).parent_path();
            candidates.push_back(dir / (python ? "polishing/worker.py" : "hprqp_polish_worker"));
            for (int up = 0; up < 3 && !dir.empty(); ++up, dir = dir.parent_path())
                candidates.push_back(dir / (python ? "share/hprqp/polishing/worker.py" : "libexec/hprqp/hprqp_polish_worker"));
        }
    }
    for (const auto &c : candidates) {
        std::error_code ec;
        if (fs::is_regular_file(c, ec) && (python || access(c.c_str(), X_OK) == 0)) {
            const fs::path abs = fs::absolute(c, ec);
            if (!ec) return abs.string();
        }
    }
    return std::string();
}

// worker_path.hpp:42-52: this library's file and the running executable
std::vector<std::string> default_anchors()
{
    std::vector<std::string> out;
    Dl_info info{};
    if (dladdr(reinterpret_cast<const void *>(&default_anchors), &info) && info.dli_fname)
        out.push_back(fs::absolute(info.dli_fname).string());
    std::error_code ec;
    const fs::path exe = fs::read_symlink("/proc/self/exe", ec);
    if (!ec) out.push_back(exe.string());
    return out;
}

// ------------------------------------------------------------------- owner
namespace {
// include/shared/diagnostics.h: feature diagnostics go to stderr only
// when HPRQP_DIAGNOSTICS=1
bool diagnostics_on()
{
    const char *v = std::getenv("HPRQP_DIAGNOSTICS");
    return v && std::strcmp(v, "1") == 0;
}
}  // namespace

// polish.cpp:66-72
Owner::Owner(const HPRQPModel *model, const HPRQPParams *params) : params_(*params), model_(model)
{
    python_ = params->polish_python ? params->polish_python : "";
    directory_ = params->polish_directory ? params->polish_directory : "/tmp";
    worker_ = find_worker(params->polish_worker, default_anchors(), !python_.empty());
    if (worker_.empty()) throw std::runtime_error("CPU worker unavailable; build/install hprqp_polish_worker");
    if (params->polish_episode_seconds < 0 || params->polish_total_seconds < 0 ||
        !std::isfinite(params->polish_episode_seconds) || !std::isfinite(params->polish_total_seconds))
        throw std::runtime_error("invalid worker budgets");
    // the strings are owned here; the copied pointers are not used again
    params_.polish_python = params_.polish_worker = params_.polish_directory = nullptr;
}

// polish.cpp:140-143
Owner *Owner::create(const HPRQPModel *model, const HPRQPParams *params)
{
    if (!model || !params || !params->polishing || model->num_var_cones) return nullptr;
    // The usual case, no worker installed, is settled without an exception:
    // a first throw in a process with many large libraries costs tens of ms.
    if (find_worker(params->polish_worker, default_anchors(), params->polish_python != nullptr).empty()) {
        if (params->verbose && diagnostics_on())
            std::fprintf(stderr, "POLISH_DISABLED CPU worker unavailable; build/install hprqp_polish_worker\n");
        return nullptr;
    }
    try {
        return new Owner(model, params);
    } catch (const std::exception &e) {
        if (params->verbose && diagnostics_on()) std::fprintf(stderr, "POLISH_DISABLED %s\n", e.what());
        return nullptr;
    }
}

// polish.cpp:73-76
Owner::~Owner()
{
    child_.cancel();
    if (!run_.empty() && !params_.polish_keep_files) {
        std::error_code ec;
        fs::remove_all(run_, ec);
    }
}

// polish.cpp:77-81
void Owner::event(const char *what, size_t iter)
{
    if (run_.empty()) return;
    std::ofstream f(run_ + "/events.jsonl", std::ios::app);
    f << "{\"event\":" << json_string(what) << ",\"iteration\":" << iter << ",\"attempt\":" << attempts_ << "}\n";
    if (params_.verbose && diagnostics_on()) std::fprintf(stderr, "POLISH %s iteration=%zu attempt=%zu\n", what, iter, attempts_);
}

// polish.cpp:145-151: any failure disables polishing for this solve
bool Owner::poll(Backend *backend, const HPRQPResiduals &base, size_t iter, double seconds, double *x, double *w,
                 double *z, double *y, HPRQPResiduals *out)
{
    if (disabled_) return false;
    try {
        return poll_impl(backend, base, iter, seconds, x, w, z, y, out);
    } catch (const std::exception &e) {
        child_.cancel();
        pending_ = false;
        disabled_ = true;
        event("DISABLED", iter);
        if (params_.verbose && diagnostics_on()) std::fprintf(stderr, "POLISH_DISABLED %s\n", e.what());
        return false;
    }
}

// polish.cpp:82-133
bool Owner::poll_impl(Backend *backend, const HPRQPResiduals &base, size_t iter, double seconds, double *x, double *w,
                      double *z, double *y, HPRQPResiduals *out)
{
    const auto tick = Clock::now();
    auto remaining = [&] {
        return params_.time_limit - seconds - std::chrono::duration<double>(Clock::now() - tick).count();
    };
    if (remaining() <= 0) {
        child_.cancel();
        pending_ = false;
        return false;
    }
    const size_t n = model_->Q.ncol, m = model_->A.nrow;
    if (pending_) {
        const Worker::State state = child_.poll();
        if (state == Worker::RUNNING) return false;
        pending_ = false;
        child_seconds_ += child_.age();
        if (state != Worker::EXITED || child_.exit_code() != 0 || !ready_matches(episode_ + "/READY", token_)) {
            event("NO_VALID_PUBLICATION", iter);
            return false;
        }
        const Certificate cert = read_certificate(episode_, token_, n, m, params_.stoptol);
        const Candidate c = read_candidate(episode_ + "/candidate.bin", n, m);
        if (remaining() <= 0) return false;
        std::copy(c.x.begin(), c.x.end(), x);
        std::copy(c.w.begin(), c.w.end(), w);
        std::copy(c.z.begin(), c.z.end(), z);
        std::copy(c.y.begin(), c.y.end(), y);
        *out = HPRQPResiduals{};
        out->is_updated = 1;
        out->err_Rp_org_bar = cert.primal_feasibility;
        out->err_Rd_org_bar = cert.dual_feasibility;
        out->rel_gap_bar = cert.gap;
        out->primal_obj_bar = cert.primal_objective;
        out->dual_obj_bar = cert.dual_objective;
        out->KKTx_and_gap_org_bar = std::max({cert.primal_feasibility, cert.dual_feasibility, cert.gap});
        accepted_ = true;
        event("CPU_CERTIFICATE_ACCEPT", iter);
        return true;
    }
    if ((params_.polish_max_attempts && attempts_ >= params_.polish_max_attempts) ||
        (params_.polish_total_seconds > 0 && child_seconds_ >= params_.polish_total_seconds) ||
        !progress_.eligible(iter, params_.check_iter, base.err_Rp_org_bar, base.err_Rd_org_bar,
                            base.KKTx_and_gap_org_bar, params_.stoptol))
        return false;
    if (run_.empty()) {
        std::string pattern = (fs::absolute(directory_) / "hprqp-polish-XXXXXX").string();
        std::vector<char> buf(pattern.begin(), pattern.end());
        buf.push_back('\0');
        if (!mkdtemp(buf.data())) throw std::runtime_error("private worker directory");
        run_ = buf.data();
        write_model(*model_, run_ + "/model.bin");
    }
    progress_.launched(base.err_Rp_org_bar, base.err_Rd_org_bar, base.KKTx_and_gap_org_bar, params_.stoptol);
    ++attempts_;
    token_ = new_token();
    episode_ = run_ + "/episode_" + std::to_string(attempts_);
    if (!fs::create_directory(episode_)) throw std::runtime_error("episode collision");
    std::vector<double> sx(n), sw(n), sz(n), sy(m);
    if (backend->copy_output_iterates(sx.data(), sw.data(), sz.data(), sy.data()) != HPRQP_STATUS_OK)
        throw std::runtime_error("snapshot export failed");
    write_snapshot(episode_ + "/snapshot.bin", n, m, sx.data(), sy.data(), sz.data(), sw.data());
    double budget = remaining();
    if (params_.polish_episode_seconds > 0) budget = std::min(budget, params_.polish_episode_seconds);
    if (params_.polish_total_seconds > 0) budget = std::min(budget, params_.polish_total_seconds - child_seconds_);
    if (!(budget > 0) || !std::isfinite(budget)) return false;
    {
        std::ofstream req(episode_ + "/request.json");
        req.exceptions(std::ios::badbit | std::ios::failbit);
        req.precision(17);
        req << "{\"token\":" << json_string(token_) << ",\"parent_pid\":" << getpid()
            << ",\"model\":" << json_string(run_ + "/model.bin") << ",\"dimensions\":[" << n << "," << m
            << "],\"iteration\":" << iter << ",\"tolerance\":" << params_.stoptol << ",\"seconds\":" << budget
            << ",\"memory_bytes\":" << params_.polish_memory_bytes << ",\"config\":{\"backend\":\"auto\"}}\n";
    }
    std::vector<std::string> command{worker_, episode_ + "/request.json"};
    if (!python_.empty()) command.insert(command.begin(), python_);
    child_.start(command, episode_ + "/worker.log", budget);
    pending_ = true;
    event("LAUNCHED", iter);
    return false;
}

}  // namespace hprqp::polish
