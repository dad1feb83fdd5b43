// CPU checks of the host polishing owner (hpr-qp-c_amd/csrc/polish_owner.cpp),
// built and run by tests/test_polish.py:
//   * the launch rule against the cases of the reference's
//     tests/polishing/test_progress.cpp;
//   * worker processes (time-out, cancel, exit status, pinned thread counts,
//     the whole process group stopped) and the publication / certificate
//     checks of tests/polishing/test_publication.cpp;
//   * worker lookup as tests/polishing/test_worker_path.cpp (without the
//     reference's source-tree fallback, which has no counterpart here; "not
//     found" is an empty path rather than an exception);
//   * SHA-256 on the FIPS 180-4 example messages.
// Usage: polish_driver <scratch dir> [file to hash]
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <string>
#include <thread>
#include <unistd.h>

#include "polish_owner.h"

namespace fs = std::filesystem;
using namespace hprqp::polish;

static int failures = 0;
#define CHECK(cond)                                                                    \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::fprintf(stderr, "FAIL line %d: %s\n", __LINE__, #cond);               \
            ++failures;                                                                \
        }                                                                              \
    } while (0)

static bool throws(const std::function<void()> &f)
{
    try {
        f();
    } catch (const std::exception &) {
        return true;
    }
    return false;
}

static std::string hexdigest(const std::array<unsigned char, 32> &d)
{
    static const char *h = "0123456789abcdef";
    std::string s;
    for (unsigned char b : d) {
        s += h[b >> 4];
        s += h[b & 15];
    }
    return s;
}

static void wait_exit(Worker &w)
{
    for (int i = 0; i < 5000 && w.poll() == Worker::RUNNING; ++i) std::this_thread::sleep_for(std::chrono::milliseconds(2));
}

static void progress_cases()
{
    Milestones p;
    CHECK(!p.eligible(0, 100, 1e-4, 1e-4, 1e-3, 1e-8));  // before 7 * check_iter
    CHECK(p.eligible(700, 100, 1e-3, 1e-3, .5, 1e-8));
    p.launched(1e-3, 1e-3, .5, 1e-8);
    CHECK(!p.eligible(800, 100, .1, .1, .5, 1e-8));
    CHECK(p.eligible(900, 100, 1e-5, 1e-5, .5, 1e-8));
    CHECK(p.eligible(1000, 100, 1e-6, 1e-6, .5, 1e-8));  // crossings while busy are not consumed
    p.launched(1e-6, 1e-6, .5, 1e-8);
    CHECK(!p.eligible(1100, 100, 1e-5, 1e-5, .5, 1e-8));
    CHECK(p.eligible(1200, 100, 1e-7, 1e-7, .5, 1e-8));
    CHECK(!p.eligible(1300, 100, 1e-9, 1e-9, 1e-9, 1e-8));  // already within tolerance
    p.launched(1e-7, 1e-7, .5, 1e-8);
    CHECK(p.eligible(1400, 100, 1e-8, 1e-8, .5, 1e-8));
    p.launched(1e-8, 1e-8, .5, 1e-8);
    CHECK(!p.eligible(1500, 100, 1e-8, 1e-8, .5, 1e-8));
    CHECK(p.eligible(1600, 100, 1e-12, 1e-12, .5, 1e-8));
    p.launched(0., 0., .5, 1e-8);
    CHECK(p.equation_target == 0.);
    CHECK(!p.eligible(1700, 100, 0., 0., .5, 1e-8));
    CHECK(!p.eligible(1800, 100, -1., 0., .5, 1e-8));
    CHECK(!p.eligible(1900, 100, 0., 0., NAN, 1e-8));
    CHECK(p.launches == 5);
    // the total target: decades of the KKT error down to 10 tol
    Milestones q;
    CHECK(q.eligible(700, 100, 1.0, 1.0, 5e-3, 1e-6));
    q.launched(1.0, 1.0, 5e-3, 1e-6);
    CHECK(!q.eligible(800, 100, 1.0, 1.0, 2e-3, 1e-6));
    CHECK(q.eligible(800, 100, 1.0, 1.0, 4e-4, 1e-6));
    q.launched(1.0, 1.0, 1e-6 * 1.5, 1e-6);  // crossing the floor: the target ends at 0
    CHECK(q.total_target == 0.);
    CHECK(!q.eligible(900, 100, 1.0, 1.0, 1.2e-6, 1e-6));
    CHECK(!q.eligible(6, 1, 1e-3, 1e-3, 1.0, 1e-6));
    Milestones r;
    CHECK(r.eligible(7, 0, 1e-3, 1e-3, 1.0, 1e-6));  // check_iter 0 counts as 1
    CHECK(!r.eligible(700, 100, 1e-3, 1e-3, 1.0, 0.0));  // no positive tolerance
}

static void worker_cases(const std::string &dir)
{
    Worker w;
    w.start({"/bin/sleep", "10"}, dir + "/timeout.log", 0.01);
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    CHECK(w.poll() == Worker::TIMED_OUT);
    CHECK(!w.running());
    w.start({"/bin/sleep", "10"}, dir + "/cancel.log", 10);
    w.cancel();
    CHECK(w.state() == Worker::CANCELLED);
    w.start({"/bin/false"}, dir + "/crash.log", 10);
    wait_exit(w);
    CHECK(w.state() == Worker::EXITED && w.exit_code() == 1);
    setenv("OPENBLAS_NUM_THREADS", "7", 1);
    w.start({"/bin/sh", "-c", "test \"$OPENBLAS_NUM_THREADS\" = 1 && test \"$OMP_NUM_THREADS\" = 1"},
            dir + "/threads.log", 10);
    wait_exit(w);
    CHECK(w.exit_code() == 0);
    CHECK(std::string(std::getenv("OPENBLAS_NUM_THREADS")) == "7");
    // the log is created fresh: an existing file refuses the start
    CHECK(throws([&] { w.start({"/bin/true"}, dir + "/threads.log", 10); }));
    CHECK(!w.running());
    // stopping a worker stops the processes it started (its group)
    const std::string pidfile = dir + "/grandchild.pid";
    w.start({"/bin/sh", "-c", "sleep 30 & echo $! > " + pidfile + "; wait"}, dir + "/group.log", 10);
    for (int i = 0; i < 2500 && !fs::exists(pidfile); ++i) std::this_thread::sleep_for(std::chrono::milliseconds(2));
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    int grandchild = 0;
    std::ifstream(pidfile) >> grandchild;
    CHECK(grandchild > 0);
    w.cancel();
    bool gone = false;
    for (int i = 0; i < 500 && !gone; ++i) {  // killed: gone, or a zombie until init reaps it
        if (grandchild <= 0) break;
        std::string stat;
        std::getline(std::ifstream("/proc/" + std::to_string(grandchild) + "/stat"), stat);
        const size_t close = stat.rfind(')');
        gone = kill(grandchild, 0) != 0 || stat.empty() || (close != std::string::npos && close + 2 < stat.size() &&
                                                             stat[close + 2] == 'Z');
        if (!gone) std::this_thread::sleep_for(std::chrono::milliseconds(2));
    }
    CHECK(gone);
    CHECK(throws([&] { w.start({}, dir + "/empty.log", 10); }));
    CHECK(throws([&] { w.start({"/bin/true"}, dir + "/budget.log", 0.0); }));
}

static void publication_cases(const std::string &ep)
{
    const std::string token(32, '0');
    {
        std::ofstream f(ep + "/READY");
        f << "HPRCPU01\n" << token << "\n";
    }
    CHECK(ready_matches(ep + "/READY", token));
    CHECK(!ready_matches(ep + "/READY", std::string(32, '1')));
    {
        std::ofstream f(ep + "/READY", std::ios::app);
        f << "extra";
    }
    CHECK(!ready_matches(ep + "/READY", token));
    CHECK(!ready_matches(ep + "/missing", token));
    auto candidate = [&](double x0) {
        std::ofstream f(ep + "/candidate.bin", std::ios::binary);
        f.write("HPRPOL01", 8);
        const uint64_t dims[2] = {1, 0};
        f.write(reinterpret_cast<const char *>(dims), 16);
        const double data[3] = {x0, 2.0, 3.0};
        f.write(reinterpret_cast<const char *>(data), 24);
    };
    auto certificate = [&](double err, double tol, double pobj = 0.0, double dobj = 0.0, double gap = 0.0) {
        std::ofstream f(ep + "/certificate.bin", std::ios::binary);
        f.write("HPRCPU01", 8);
        const char id[16] = {};
        f.write(id, 16);
        const uint64_t dims[2] = {1, 0};
        f.write(reinterpret_cast<const char *>(dims), 16);
        const auto d = sha256_file(ep + "/candidate.bin");
        f.write(reinterpret_cast<const char *>(d.data()), 32);
        const double v[6] = {tol, err, 0.0, gap, pobj, dobj};
        f.write(reinterpret_cast<const char *>(v), 48);
    };
    candidate(1.0);
    certificate(0.0, 1e-8);
    CHECK(!throws([&] { read_certificate(ep, token, 1, 0, 1e-8); }));
    const Candidate c = read_candidate(ep + "/candidate.bin", 1, 0);
    CHECK(c.x.size() == 1 && c.x[0] == 1.0 && c.w[0] == 2.0 && c.z[0] == 3.0 && c.y.empty() && c.s.empty());
    CHECK(throws([&] { read_candidate(ep + "/candidate.bin", 2, 0); }));
    CHECK(throws([&] { read_certificate(ep, std::string(32, '1'), 1, 0, 1e-8); }));  // another request
    CHECK(throws([&] { read_certificate(ep, token, 2, 0, 1e-8); }));                 // other dimensions
    {
        std::fstream f(ep + "/candidate.bin", std::ios::in | std::ios::out | std::ios::binary);
        f.seekp(24);
        const double x = 5.0;
        f.write(reinterpret_cast<const char *>(&x), 8);
    }
    CHECK(throws([&] { read_certificate(ep, token, 1, 0, 1e-8); }));  // payload changed after publication
    certificate(1e-4, 1e-8);
    CHECK(throws([&] { read_certificate(ep, token, 1, 0, 1e-8); }));  // not accurate enough
    certificate(NAN, 1e-8);
    CHECK(throws([&] { read_certificate(ep, token, 1, 0, 1e-8); }));
    certificate(0.0, 1e-6);
    CHECK(throws([&] { read_certificate(ep, token, 1, 0, 1e-8); }));  // another tolerance
    certificate(0.0, 1e-8, 1.0, 2.0, 1.0 / 3.0);
    CHECK(throws([&] { read_certificate(ep, token, 1, 0, 1e-8); }));  // objectives disagree
    certificate(0.0, 1e-8, 1.0, 1.0, 1e-9);
    CHECK(throws([&] { read_certificate(ep, token, 1, 0, 1e-8); }));  // recorded gap inconsistent
    certificate(0.0, 1e-8, 1.0, 1.0, 0.0);
    CHECK(!throws([&] { read_certificate(ep, token, 1, 0, 1e-8); }));
    {
        std::ofstream f(ep + "/candidate.bin", std::ios::binary | std::ios::app);
        f.write("x", 1);
    }
    CHECK(throws([&] { read_certificate(ep, token, 1, 0, 1e-8); }));  // wrong candidate length
}

static void lookup_cases(const std::string &root)
{
    const fs::path prefix = fs::path(root) / "moved";
    auto put = [](const fs::path &p) {
        fs::create_directories(p.parent_path());
        std::ofstream(p) << "# test\n";
        fs::permissions(p, fs::perms::owner_exec, fs::perm_options::add);
    };
    const fs::path installed = prefix / "libexec/hprqp/hprqp_polish_worker";
    put(installed);
    CHECK(find_worker(nullptr, {(prefix / "bin/solve_qp").string()}, false) == installed.string());
    CHECK(find_worker(nullptr, {(prefix / "lib/libhprqp.so").string()}, false) == installed.string());
    CHECK(find_worker(nullptr, {(prefix / "lib/x86_64-linux-gnu/libhprqp.so").string()}, false) == installed.string());
    const fs::path other = fs::path(root) / "elsewhere/worker";
    put(other);
    CHECK(find_worker(other.c_str(), {(prefix / "bin/solve_qp").string()}, false) == other.string());
    const fs::path adjacent = prefix / "bin/hprqp_polish_worker";
    put(adjacent);
    CHECK(find_worker(nullptr, {(prefix / "bin/solve_qp").string()}, false) == adjacent.string());
    fs::remove(adjacent);
    fs::permissions(installed, fs::perms::owner_exec, fs::perm_options::remove);
    CHECK(find_worker(nullptr, {(prefix / "bin/solve_qp").string()}, false).empty());  // not executable
    CHECK(find_worker("/not/a/worker", {}, false).empty());
    CHECK(find_worker(nullptr, {}, false).empty());
    const fs::path script = prefix / "share/hprqp/polishing/worker.py";
    fs::create_directories(script.parent_path());
    std::ofstream(script) << "# test\n";  // a Python worker need not be executable
    CHECK(find_worker(nullptr, {(prefix / "bin/solve_qp").string()}, true) == script.string());
    const auto anchors = default_anchors();
    CHECK(!anchors.empty());
}

static void sha_cases()
{
    CHECK(hexdigest(sha256("", 0)) == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855");
    CHECK(hexdigest(sha256("abc", 3)) == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad");
    const char *two = "abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq";
    CHECK(hexdigest(sha256(two, std::strlen(two))) == "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1");
    const std::string million(1000000, 'a');
    CHECK(hexdigest(sha256(million.data(), million.size())) ==
          "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0");
}

// A 2-variable, 1-row model and a snapshot, written for the Python side of
// the test to parse.
static void write_files(const std::string &dir)
{
    size_t qrp[3] = {0, 2, 3}, qci[3] = {0, 1, 1};
    double qv[3] = {2.0, 0.5, 3.0};
    size_t arp[2] = {0, 2}, aci[2] = {0, 1};
    double av[2] = {1.0, -1.0};
    double c[2] = {1.0, -2.0}, l[2] = {-1.0, -INFINITY}, u[2] = {4.0, INFINITY}, AL[1] = {0.5}, AU[1] = {2.5};
    HPRQPModel md{};
    md.Q = HPRQPMatrixCSR{2, 2, 3, qrp, qci, qv, 0};
    md.A = HPRQPMatrixCSR{1, 2, 2, arp, aci, av, 0};
    md.c = c;
    md.l = l;
    md.u = u;
    md.AL = AL;
    md.AU = AU;
    md.obj_constant = 7.25;
    write_model(md, dir + "/model.bin");
    const double x[2] = {1.5, -0.5}, y[1] = {0.25}, z[2] = {0.0, 1.0}, w[2] = {3.0, 4.0};
    write_snapshot(dir + "/snapshot.bin", 2, 1, x, y, z, w);
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    progress_cases();
    worker_cases(dir);
    fs::create_directories(dir + "/episode");
    publication_cases(dir + "/episode");
    lookup_cases(dir + "/lookup");
    sha_cases();
    write_files(dir);
    if (argc > 2) std::printf("sha256 %s\n", hexdigest(sha256_file(argv[2])).c_str());
    std::printf("failures %d\n", failures);
    return failures ? 1 : 0;
}
