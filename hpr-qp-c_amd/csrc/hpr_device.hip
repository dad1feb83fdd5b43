// MI355X (gfx950) backend of the HPR-QP iteration.
//
// Replaces the reference's CUDA backend (src/gpu/*.cu) behind the same host
// loop contract. Layout in HBM: int32 CSR for Q, A and the device-built A^T
// (plus the permutation that maps A^T and Q^T entries to their sources), FP64
// values, FP64 vectors in one arena. Every per-iteration kernel is a row-blocked
// SpMV whose per-row result is consumed by a fused epilogue (the HPR update for
// that row); reductions are two-level and deterministic, and one pinned
// read-back per residual / M-norm check is the only host synchronisation.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "device_common.h"
#include "epilogues.h"
#include "gram.h"
#include "hprqp.h"
#include "hprqp_hip.h"
#include "solve_loop.h"

namespace hprqp {

void report_hip_error(hipError_t err, const char *what, const char *file, int line)
{
    std::fprintf(stderr, "hprqp: HIP error %d (%s) in %s at %s:%d\n", (int)err, hipGetErrorString(err), what, file,
                 line);
}

__global__ __launch_bounds__(kThreads) void reduce_partials_kernel(const double *__restrict__ partials, int nblk,
                                                                   int stride, int ns, double *__restrict__ out)
{
    const int term = blockIdx.x;
    const bool is_max = term >= ns;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kThreads) {
        double v = partials[(size_t)b * stride + term];
        acc = is_max ? fmax(acc, v) : acc + v;
    }
    __shared__ double part[kThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    acc = is_max ? wave_max(acc) : wave_sum(acc);
    if (lane == 0) part[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double total = part[0];
        for (int w = 1; w < kThreads / kWave; ++w) total = is_max ? fmax(total, part[w]) : total + part[w];
        out[term] = total;
    }
}

// Several finished reductions in one launch: block b reduces term b -
// first[j] of job j (the same arithmetic as reduce_partials_kernel, so the
// same bits); up to kMaxRedJobs jobs, their terms numbered consecutively.
constexpr int kMaxRedJobs = 4;
struct RedJobs {
    const double *partials[kMaxRedJobs];
    double *out[kMaxRedJobs];
    int nblk[kMaxRedJobs], nterms[kMaxRedJobs], ns[kMaxRedJobs], first[kMaxRedJobs];
    int count;
};

__global__ __launch_bounds__(kThreads) void reduce_jobs_kernel(RedJobs jobs)
{
    int j = 0;
    while (j + 1 < jobs.count && (int)blockIdx.x >= jobs.first[j + 1]) ++j;
    const int term = (int)blockIdx.x - jobs.first[j];
    const int stride = jobs.nterms[j];
    const bool is_max = term >= jobs.ns[j];
    double acc = 0.0;
    for (int b = threadIdx.x; b < jobs.nblk[j]; b += kThreads) {
        double v = jobs.partials[j][(size_t)b * stride + term];
        acc = is_max ? fmax(acc, v) : acc + v;
    }
    __shared__ double part[kThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    acc = is_max ? wave_max(acc) : wave_sum(acc);
    if (lane == 0) part[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double total = part[0];
        for (int w = 1; w < kThreads / kWave; ++w) total = is_max ? fmax(total, part[w]) : total + part[w];
        jobs.out[j][term] = total;
    }
}

namespace {

inline int grid_for(size_t len) { return (int)((len + kThreads - 1) / kThreads); }

// red_ holds the reduction results: indices below kPowSlots for the residuals,
// the sigma update and the setup; from kPowSlots one slot of three per power
// step of a run.
constexpr int kPowRun = 16;  // power-iteration steps issued per host read
constexpr int kPowSlots = 64;
constexpr int kRedLen = kPowSlots + 3 * kPowRun;

// HPRQP_SETUP_TRACE=1 reports the wall time of each setup stage on stderr,
// waiting for the stream at each mark (diagnostics only).
struct SetupClock {
    hipStream_t s;
    bool on = std::getenv("HPRQP_SETUP_TRACE") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    explicit SetupClock(hipStream_t stream) : s(stream) {}
    void mark(const char *what)
    {
        if (!on) return;
        (void)hipStreamSynchronize(s);
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "hprqp setup: %-34s %9.3f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

// ------------------------------------------------------------------ device CSR

struct DevCsr {
    int nrow = 0, ncol = 0, nnz = 0;
    int *rp = nullptr;
    int *ci = nullptr;
    double *val = nullptr;
    int2 *blocks = nullptr;  // {first row, first entry} per row block, nblocks + 1 entries
    int nblocks = 0;
};

}  // namespace

// Host planning of row blocks from the row pointer (see spmv_rows_kernel):
// runs of rows of <= kLongRow entries become stream blocks (<= kThreads rows,
// <= kStreamCap entries), runs of longer rows become wave blocks of up to
// kThreads / kWave rows, and a row longer than kVecRow a block of its own.
// Entry b is {first row, first entry} of block b; the last entry closes it.
// `rp(i)` returns row_ptr[i] and is called with nondecreasing i, except that
// after rp(i + 1) the next call may be rp(i); `emit` receives the entries in
// order.
template <class RowPtr, class Emit>
void plan_row_blocks_with(int nrow, RowPtr &&rp, Emit &&emit)
{
    int r = 0;
    while (r < nrow) {
        const int start = rp(r);
        const int len = rp(r + 1) - start;
        emit(make_int2(r, start));
        const int first = r;
        if (len > kVecRow) {
            ++r;
        } else if (len > kLongRow) {
            while (r < nrow && r - first < kThreads / kWave) {
                const int b = rp(r);
                const int l2 = rp(r + 1) - b;
                if (l2 <= kLongRow || l2 > kVecRow) break;
                ++r;
            }
        } else {
            int nnz = 0;
            while (r < nrow && r - first < kThreads) {
                const int b = rp(r);
                const int l2 = rp(r + 1) - b;
                if (l2 > kLongRow || nnz + l2 > kStreamCap) break;
                nnz += l2;
                ++r;
            }
        }
    }
    emit(make_int2(nrow, nrow >= 0 ? rp(nrow) : 0));
}

std::vector<int2> plan_row_blocks(const std::vector<int> &rp)
{
    const int nrow = (int)rp.size() - 1;
    if (nrow == 0) return {make_int2(0, 0), make_int2(0, 0)};
    std::vector<int2> blk;
    blk.reserve(nrow / 16 + 2);
    plan_row_blocks_with(
        nrow, [&](int i) { return rp[i]; }, [&](int2 e) { blk.push_back(e); });
    return blk;
}

namespace {

HPRQPStatus upload_blocks(DevCsr &m, hipStream_t s)
{
    std::vector<int> rp(m.nrow + 1);
    HPRQP_HIP(hipMemcpyAsync(rp.data(), m.rp, sizeof(int) * (m.nrow + 1), hipMemcpyDeviceToHost, s));
    HPRQP_HIP(hipStreamSynchronize(s));
    std::vector<int2> blk = plan_row_blocks(rp);
    m.nblocks = (int)blk.size() - 1;
    if (m.blocks) HPRQP_HIP(hipFree(m.blocks));
    HPRQP_HIP(hipMalloc(&m.blocks, sizeof(int2) * blk.size()));
    HPRQP_HIP(hipMemcpyAsync(m.blocks, blk.data(), sizeof(int2) * blk.size(), hipMemcpyHostToDevice, s));
    HPRQP_HIP(hipStreamSynchronize(s));
    return HPRQP_STATUS_OK;
}

// Column panels of one matrix (see spmv_rows_kernel): part[p] holds the
// entries whose column lies in [p*width, (p+1)*width), rows unchanged;
// src[p][k] is the entry's index in the full matrix (values are gathered
// through it after scaling).
constexpr int kMaxPanels = 4;
struct Panels {
    int P = 1;
    int width = 0;
    unsigned lds_pad = 0;  // unused dynamic LDS per workgroup: caps residency (autotune_panels)
    DevCsr part[kMaxPanels];
    int *src[kMaxPanels] = {nullptr, nullptr, nullptr, nullptr};
};

// Destinations of the device-side panel split (one entry per panel).
struct PanelOut {
    const int *rp[kMaxPanels];
    int *ci[kMaxPanels];
    int *src[kMaxPanels];
};

void free_csr(DevCsr &m);
HPRQPStatus upload_blocks(DevCsr &m, hipStream_t s);

void free_panels(Panels &pn)
{
    for (int p = 0; p < kMaxPanels; ++p) {
        free_csr(pn.part[p]);
        if (pn.src[p]) (void)hipFree(pn.src[p]);
        pn.src[p] = nullptr;
    }
    pn.P = 1;
    pn.width = 0;
}

void free_csr(DevCsr &m)
{
    if (m.rp) (void)hipFree(m.rp);
    if (m.ci) (void)hipFree(m.ci);
    if (m.val) (void)hipFree(m.val);
    if (m.blocks) (void)hipFree(m.blocks);
    m = DevCsr{};
}

__global__ void gather_values_kernel(int nnz, const int *__restrict__ perm, const double *__restrict__ src,
                                     double *__restrict__ dst);

// Column-panel split on the device. Per row (one thread each): the number
// of its entries in each panel; an inclusive scan of each panel's counts
// gives the panel's row pointers; then each row copies its entries to their
// panels in row order, recording every entry's source index.
__global__ void panel_count_kernel(int nrow, const int *__restrict__ rp, const int *__restrict__ ci, int width,
                                   int P, int *__restrict__ counts)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    int c[kMaxPanels] = {0, 0, 0, 0};
    for (int k = rp[r]; k < rp[r + 1]; ++k) ++c[min(ci[k] / width, P - 1)];
    for (int p = 0; p < P; ++p) counts[(size_t)p * (nrow + 1) + r + 1] = c[p];
    if (r == 0)
        for (int p = 0; p < P; ++p) counts[(size_t)p * (nrow + 1)] = 0;
}

__global__ void panel_scatter_kernel(int nrow, const int *__restrict__ rp, const int *__restrict__ ci, int width,
                                     int P, PanelOut out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    int pos[kMaxPanels];
    for (int p = 0; p < P; ++p) pos[p] = out.rp[p][r];
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
        const int p = min(ci[k] / width, P - 1);
        const int d = pos[p]++;
        out.ci[p][d] = ci[k];
        out.src[p][d] = k;
    }
}

// Split M into P column panels (structure built on the device from M's
// indices; values filled by fill_panel_values once they are final).
HPRQPStatus build_panels(const DevCsr &M, int P, Panels &pn, hipStream_t s)
{
    free_panels(pn);
    pn.P = P;
    pn.width = (M.ncol + P - 1) / P;
    if (P <= 1) return HPRQP_STATUS_OK;
    const size_t rows1 = (size_t)M.nrow + 1;
    int *counts = nullptr;
    HPRQP_HIP(hipMalloc(&counts, sizeof(int) * rows1 * P));
    if (M.nrow > 0)
        panel_count_kernel<<<grid_for(M.nrow), kThreads, 0, s>>>(M.nrow, M.rp, M.ci, pn.width, P, counts);
    else
        HPRQP_HIP(hipMemsetAsync(counts, 0, sizeof(int) * rows1 * P, s));
    HPRQP_HIP(hipGetLastError());
    PanelOut out{};
    std::vector<int> nnz(P, 0);
    for (int p = 0; p < P; ++p) {
        DevCsr &d = pn.part[p];
        d.nrow = M.nrow;
        d.ncol = M.ncol;
        HPRQP_HIP(hipMalloc(&d.rp, sizeof(int) * rows1));
        size_t tmp_bytes = 0;
        HPRQP_HIP(rocprim::inclusive_scan(nullptr, tmp_bytes, counts + p * rows1, d.rp, rows1, rocprim::plus<int>(), s));
        void *tmp = nullptr;
        HPRQP_HIP(hipMalloc(&tmp, std::max(tmp_bytes, (size_t)4)));
        HPRQP_HIP(rocprim::inclusive_scan(tmp, tmp_bytes, counts + p * rows1, d.rp, rows1, rocprim::plus<int>(), s));
        HPRQP_HIP(hipMemcpyAsync(&nnz[p], d.rp + M.nrow, sizeof(int), hipMemcpyDeviceToHost, s));
        HPRQP_HIP(hipStreamSynchronize(s));
        (void)hipFree(tmp);
        d.nnz = nnz[p];
        const size_t nz = std::max((size_t)d.nnz, (size_t)1);
        HPRQP_HIP(hipMalloc(&d.ci, sizeof(int) * nz));
        HPRQP_HIP(hipMalloc(&d.val, sizeof(double) * nz));
        HPRQP_HIP(hipMalloc(&pn.src[p], sizeof(int) * nz));
        out.rp[p] = d.rp;
        out.ci[p] = d.ci;
        out.src[p] = pn.src[p];
    }
    if (M.nrow > 0) panel_scatter_kernel<<<grid_for(M.nrow), kThreads, 0, s>>>(M.nrow, M.rp, M.ci, pn.width, P, out);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(hipStreamSynchronize(s));
    (void)hipFree(counts);
    for (int p = 0; p < P; ++p) {
        HPRQPStatus st = upload_blocks(pn.part[p], s);
        if (st != HPRQP_STATUS_OK) return st;
    }
    return HPRQP_STATUS_OK;
}

HPRQPStatus fill_panel_values(const DevCsr &M, Panels &pn, hipStream_t s)
{
    if (pn.P <= 1) return HPRQP_STATUS_OK;
    for (int p = 0; p < pn.P; ++p)
        if (pn.part[p].nnz)
            gather_values_kernel<<<grid_for(pn.part[p].nnz), kThreads, 0, s>>>(pn.part[p].nnz, pn.src[p], M.val,
                                                                              pn.part[p].val);
    HPRQP_HIP(hipGetLastError());
    return HPRQP_STATUS_OK;
}

// ----------------------------------------------------------- device transpose
// Stable counting-sort transpose: entries are sorted by column with an LSD
// radix sort keyed on the column (stable), carrying the source entry index.
// Source entries are in row-major order, so within each output row the source
// rows come out ascending, which is the ordering of the host transpose.

__global__ void expand_rows_kernel(int nrow, const int *__restrict__ rp, int *__restrict__ row_of)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    for (int k = rp[r]; k < rp[r + 1]; ++k) row_of[k] = r;
}

__global__ void iota_kernel(int n, int *__restrict__ v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

// t_row_ptr from sorted keys: rp[j] = first position with key >= j.
__global__ void row_ptr_from_sorted_kernel(int nnz, int ncol, const int *__restrict__ keys, int *__restrict__ trp)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > nnz) return;
    int prev = p == 0 ? -1 : keys[p - 1];
    int cur = p == nnz ? ncol : keys[p];
    for (int j = prev + 1; j <= cur; ++j) trp[j] = p;
}

__global__ void gather_transpose_kernel(int nnz, const int *__restrict__ perm, const int *__restrict__ row_of,
                                        const double *__restrict__ val, int *__restrict__ tci,
                                        double *__restrict__ tval)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    int k = perm[p];
    if (tci) tci[p] = row_of[k];
    if (tval) tval[p] = val[k];
}

__global__ void gather_values_kernel(int nnz, const int *__restrict__ perm, const double *__restrict__ src,
                                     double *__restrict__ dst)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) dst[p] = src[perm[p]];
}

HPRQPStatus device_transpose(int nrow, int ncol, int nnz, const int *rp, const int *ci, const double *val, int *trp,
                             int *tci, double *tval, int *perm_out, hipStream_t s)
{
    if (nnz == 0) {
        HPRQP_HIP(hipMemsetAsync(trp, 0, sizeof(int) * (ncol + 1), s));
        return HPRQP_STATUS_OK;
    }
    int *row_of = nullptr, *keys_out = nullptr, *idx_in = nullptr, *perm = perm_out;
    bool own_perm = perm == nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    HPRQP_HIP(hipMalloc(&row_of, sizeof(int) * nnz));
    HPRQP_HIP(hipMalloc(&keys_out, sizeof(int) * nnz));
    HPRQP_HIP(hipMalloc(&idx_in, sizeof(int) * nnz));
    if (own_perm) HPRQP_HIP(hipMalloc(&perm, sizeof(int) * nnz));
    expand_rows_kernel<<<grid_for(nrow), kThreads, 0, s>>>(nrow, rp, row_of);
    iota_kernel<<<grid_for(nnz), kThreads, 0, s>>>(nnz, idx_in);
    unsigned end_bit = 1;
    while (end_bit < 31 && (1u << end_bit) < (unsigned)(ncol > 0 ? ncol : 1)) ++end_bit;
    HPRQP_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, ci, keys_out, idx_in, perm, nnz, 0, end_bit, s));
    HPRQP_HIP(hipMalloc(&tmp, tmp_bytes));
    HPRQP_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, ci, keys_out, idx_in, perm, nnz, 0, end_bit, s));
    row_ptr_from_sorted_kernel<<<grid_for(nnz + 1), kThreads, 0, s>>>(nnz, ncol, keys_out, trp);
    gather_transpose_kernel<<<grid_for(nnz), kThreads, 0, s>>>(nnz, perm, row_of, val, tci, tval);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(hipStreamSynchronize(s));
    HPRQP_HIP(hipFree(tmp));
    HPRQP_HIP(hipFree(row_of));
    HPRQP_HIP(hipFree(keys_out));
    HPRQP_HIP(hipFree(idx_in));
    if (own_perm) HPRQP_HIP(hipFree(perm));
    return HPRQP_STATUS_OK;
}

// ------------------------------------------------------------- cone (SOC)
// src/gpu/soc_common.cuh:7-34: Lorentz projection of one block (t, v):
// ||v|| by a max-scaled compensated sum; identity inside, zero in the polar
// cone, ((||v||+t)/2, (1+t/||v||)/2 v) otherwise.
__device__ void soc_project_block(double *__restrict__ out, const double *__restrict__ in, int len)
{
    const double t = in[0];
    double mx = 0.0, sum = 0.0, comp = 0.0;
    for (int j = 1; j < len; ++j) mx = fmax(mx, fabs(in[j]));
    if (mx > 0.0) {
        for (int j = 1; j < len; ++j) {
            const double sc = in[j] / mx;
            const double term = sc * sc - comp;
            const double nx = sum + term;
            comp = (nx - sum) - term;
            sum = nx;
        }
    }
    const double norm = mx * sqrt(sum);
    if (norm <= t) {
        for (int j = 0; j < len; ++j) out[j] = in[j];
    } else if (norm <= -t) {
        for (int j = 0; j < len; ++j) out[j] = 0.0;
    } else {
        const double alpha = 0.5 + 0.5 * (t / norm);
        out[0] = 0.5 * norm + 0.5 * t;
        for (int j = 1; j < len; ++j) out[j] = alpha * in[j];
    }
}

__device__ __forceinline__ double box(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

// soc_iteration.cuh:14-26 for one variable, given its projected value.
struct SocFinishX {
    double *x, *x_hat, *w_bar, *dx, *z_bar;
    const double *tempv, *w, *last_x, *fact1, *fact2;
    double sigma, f1s, f2s, a, b;
    int diag;
    __device__ void apply(int i, double xb) const
    {
        const double xi = x[i];
        const double xh = 2.0 * xb - xi;
        const double f2 = diag ? fact2[i] : f2s;
        const double f1 = diag ? fact1[i] : f1s;
        dx[i] = xb - xi;
        z_bar[i] = (xb - tempv[i]) / sigma;
        x_hat[i] = xh;
        x[i] = a * last_x[i] + b * xh;
        w_bar[i] = f1 * w[i] + f2 * xh;
    }
};

// soc_iteration.cuh:7-12 (raw x) fused with Q w; interval prefix projected
// and finished in place, cone entries left for the cone kernel.
struct EpiSocX : NoRed {
    SocFinishX fin;
    double *Qw, *tempv, *x_bar;
    const double *xv, *ATy, *c, *l, *u;
    int lin_n;
    __device__ void row(int i, double qw, double *, double *) const
    {
        Qw[i] = qw;
        const double raw = xv[i] + fin.sigma * (-qw + ATy[i] - c[i]);
        tempv[i] = raw;
        if (i < lin_n) {
            const double xb = box(raw, l[i], u[i]);
            x_bar[i] = xb;
            fin.apply(i, xb);
        }
    }
};

// Cone blocks, one thread per cone (small workgroups so that the few
// thousand cones spread over every CU): x_bar = projection of the raw point,
// the elementwise finish of the cone range then a coalesced map. The path of
// cones longer than a wavefront, or of cones that do not tile their range;
// the others take soc_cone_group_kernel below.
constexpr int kConeThreads = 64;

__global__ __launch_bounds__(kConeThreads) void soc_cone_project_kernel(int count, const int *__restrict__ off,
                                                                        const double *__restrict__ in,
                                                                        double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int b = off[k], e = off[k + 1];
    soc_project_block(out + b, in + b, e - b);
}

// soc_iteration.cuh:14-26 on the cone range [base, base + len) of x.
struct MapSocFinishX : NoRed {
    SocFinishX fin;
    const double *x_bar;
    int base;
    __device__ void row(int i, double *, double *) const { fin.apply(base + i, x_bar[base + i]); }
};

// soc_iteration.cuh:28-41 for one row
struct SocFinishY {
    double *y, *y_bar, *dy;
    const double *s, *Rp, *last_y;
    double tau, a, b;
    __device__ void apply(int i) const { apply_s(i, s[i]); }
    // the same with s_i = s[i] given
    __device__ void apply_s(int i, double si) const
    {
        const double yi = y[i];
        const double yb = (si - Rp[i]) / tau;
        dy[i] = yb - yi;
        y_bar[i] = yb;
        y[i] = a * last_y[i] + b * (2.0 * yb - yi);
    }
};

// Cone projection fused with the elementwise finish, L lanes per cone (L a
// power of two >= the longest cone of the range, <= the wavefront): lane j
// loads and writes entry j of its cone, so the accesses of a wavefront are
// contiguous. Every lane folds the squared scaled entries in entry order
// with the compensated sum of soc_project_block (the values come over
// lane shuffles), so the projection has soc_project_block's bits.
struct ConeFinX {
    SocFinishX fin;
    __device__ void operator()(int i, double xb) const { fin.apply(i, xb); }
};
struct ConeFinY {  // s[i] is the projected value this lane has just stored
    SocFinishY fin;
    __device__ void operator()(int i, double si) const { fin.apply_s(i, si); }
};

// The projection of one cone (entry j of it in lane j of an L-lane group;
// lanes past the cone's len entries give v = 0 and get back 0).
template <int L>
__device__ __forceinline__ double cone_lanes_project(double v, int j, int len)
{
    static_assert(L >= 1 && L <= kWave && (L & (L - 1)) == 0, "L lanes per cone: a power of two <= 64");
    const bool act = j < len;
    const double t = __shfl(v, 0, L);
    double mx = act && j >= 1 ? fabs(v) : 0.0;
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, L));
    const double sc = mx > 0.0 && act && j >= 1 ? v / mx : 0.0;
    double sum = 0.0, comp = 0.0;
#pragma unroll
    for (int jj = 1; jj < L; ++jj) {
        const double s_j = __shfl(sc, jj, L);
        if (mx > 0.0 && jj < len) {
            const double term = s_j * s_j - comp;
            const double nx = sum + term;
            comp = (nx - sum) - term;
            sum = nx;
        }
    }
    const double norm = mx * sqrt(sum);
    if (norm <= t) return v;
    if (norm <= -t) return 0.0;
    const double alpha = 0.5 + 0.5 * (t / norm);
    return j == 0 ? 0.5 * norm + 0.5 * t : alpha * v;
}

template <int L, class Fin>
__global__ __launch_bounds__(kThreads) void soc_cone_group_kernel(int count, const int *__restrict__ off,
                                                                  const double *__restrict__ in, double *out,
                                                                  Fin fin)
{
    const int k = (int)((blockIdx.x * (unsigned)kThreads + threadIdx.x) / L);
    const int j = (int)(threadIdx.x % L);
    int b = 0, len = 0;
    if (k < count) {
        b = off[k];
        len = off[k + 1] - b;
    }
    const bool act = j < len;
    const double xo = cone_lanes_project<L>(act ? in[b + j] : 0.0, j, len);
    if (act) {
        out[b + j] = xo;
        fin(b + j, xo);
    }
}

// The whole x step of a cone model whose Q is exactly diagonal (every row of
// Q holds at most its diagonal entry): Q w is elementwise there, so the
// interval prefix (one thread per row, the first nbox blocks) and the cones
// (L lanes per cone) each form their row's Q w and raw point themselves and
// finish in one launch, then the row's tempv for the y step (the product
// Q w_bar, elementwise too). (Q v)_i = 0 + q_ii v_i, the CSR row product's
// rounding, so the bits are those of the two products, the cone kernel and
// the finish.
template <int L>
__global__ __launch_bounds__(kThreads) void soc_x_diag_kernel(int nbox, int count, const int *__restrict__ off,
                                                              const int *__restrict__ qrp,
                                                              const double *__restrict__ qval,
                                                              const double *__restrict__ w, EpiSocX ex)
{
    auto q_times = [&](int i, double v) {
        double acc = 0.0;
        if (qrp[i] < qrp[i + 1]) acc += qval[qrp[i]] * v;
        return acc;
    };
    // iteration.c:141-143 once the finish has written x_hat_i and w_bar_i
    auto next_tempv = [&](int i, double qw) {
        const double qwb = q_times(i, ex.fin.w_bar[i]);
        ex.tempv[i] = ex.fin.x_hat[i] + ex.fin.sigma * (qw - qwb);
    };
    if ((int)blockIdx.x < nbox) {
        const int i = blockIdx.x * kThreads + threadIdx.x;
        if (i < ex.lin_n) {
            const double qw = q_times(i, w[i]);
            ex.row(i, qw, nullptr, nullptr);
            next_tempv(i, qw);
        }
        return;
    }
    const int k = (int)(((blockIdx.x - nbox) * (unsigned)kThreads + threadIdx.x) / L);
    const int j = (int)(threadIdx.x % L);
    int b = 0, len = 0;
    if (k < count) {
        b = off[k];
        len = off[k + 1] - b;
    }
    const bool act = j < len;
    double v = 0.0, qw = 0.0;
    if (act) {
        const int i = b + j;
        qw = q_times(i, w[i]);
        ex.Qw[i] = qw;
        v = ex.xv[i] + ex.fin.sigma * (-qw + ex.ATy[i] - ex.c[i]);
        ex.tempv[i] = v;
    }
    const double xo = cone_lanes_project<L>(v, j, len);
    if (act) {
        ex.x_bar[b + j] = xo;
        ex.fin.apply(b + j, xo);
        next_tempv(b + j, qw);
    }
}

struct EpiSocY : NoRed {
    SocFinishY fin;
    double *Rp, *s;
    const double *rhs, *AL, *AU;
    int lin_m;
    // equality-row metric (null when off): rows with gmap >= 0 hand
    // A t - b to the metric solve instead of finishing here
    const int *gmap = nullptr;
    double *gr = nullptr;
    const double *gb = nullptr;
    __device__ void row(int i, double ax, double *, double *) const
    {
        const double raw = ax - (rhs ? rhs[i] : 0.0) - fin.tau * fin.y[i];
        Rp[i] = raw;
        if (i < lin_m) {
            s[i] = box(raw, AL[i], AU[i]);
            const int q = gmap ? gmap[i] : -1;
            if (q >= 0) gr[q] = ax - gb[q];
            else fin.apply(i);
        }
    }
};

// soc_iteration.cuh:28-41 on the cone range [base, base + len) of y.
struct MapSocFinishY : NoRed {
    SocFinishY fin;
    int base;
    __device__ void row(int i, double *, double *) const { fin.apply(base + i); }
};

// Projection of a mixed vector (interval prefix then cones) for residuals.
__global__ void soc_project_mixed_kernel(int len, int lin, const double *__restrict__ in, const double *__restrict__ lo,
                                         const double *__restrict__ hi, int count, const int *__restrict__ off,
                                         double *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < lin) out[t] = box(in[t], lo[t], hi[t]);
    if (t < count) soc_project_block(out + off[t], in + off[t], off[t + 1] - off[t]);
}

// soc_residuals.cuh:5-9 + the dot products of :194-205 on the rows
struct EpiSocResA {
    static constexpr int NS = 2, NM = 1;
    double *partials;
    double *Ax, *arg;
    const double *rhs, *row_norm, *y_bar, *s;
    double b_scale;
    __device__ void row(int i, double ax, double *sm, double *mx) const
    {
        Ax[i] = ax;
        const double r = rhs ? rhs[i] : 0.0;
        arg[i] = ax - r;
        mx[0] = fmax(mx[0], fabs(ax * (row_norm[i] * b_scale)));
        sm[0] += y_bar[i] * s[i];
        sm[1] += y_bar[i] * r;
    }
};

// soc_residuals.cuh:11-18
struct MapSocRowResid {
    static constexpr int NS = 0, NM = 1;
    double *partials;
    const double *proj, *arg, *row_norm;
    double b_scale;
    __device__ void row(int i, double *, double *mx) const
    {
        mx[0] = fmax(mx[0], fabs((proj[i] - arg[i]) * (row_norm[i] * b_scale)));
    }
};

// soc_residuals.cuh:20-24
struct MapSocVarResid {
    static constexpr int NS = 0, NM = 1;
    double *partials;
    const double *proj, *x_bar, *col_norm;
    double b_scale;
    __device__ void row(int i, double *, double *mx) const
    {
        mx[0] = fmax(mx[0], fabs((proj[i] - x_bar[i]) * b_scale / col_norm[i]));
    }
};

// --------------------------------------------------------------- scaling
// scale.c:62-207: sqrt of column/row max (or sum) of |entries|, then divide.
// Column j of A is read through A^T's permutation, column j of Q through
// Q^T's, so that sums run in the same (row-ascending) order as the host.

template <bool SUM>
__global__ void scale_col_norm_kernel(int n, const int *__restrict__ atrp, const int *__restrict__ atperm,
                                      const double *__restrict__ aval, const int *__restrict__ qtrp,
                                      const int *__restrict__ qtperm, const double *__restrict__ qval,
                                      double *__restrict__ out)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double acc = 0.0;
    for (int k = atrp[j]; k < atrp[j + 1]; ++k) {
        double v = fabs(aval[atperm[k]]);
        if (SUM) acc += v;
        else if (v > acc) acc = v;
    }
    for (int k = qtrp[j]; k < qtrp[j + 1]; ++k) {
        double v = fabs(qval[qtperm[k]]);
        if (SUM) acc += v;
        else if (v > acc) acc = v;
    }
    acc = sqrt(acc);
    out[j] = acc == 0.0 ? 1.0 : acc;
}

// Over the matrix's row blocks (one workgroup each): rows of a stream block
// one per thread; the long rows of a wave block in turn, the max by the
// whole workgroup (order-free), the sum by one thread in entry order with
// its loads issued eight at a time (the host's sequential sum).
template <bool SUM>
__global__ __launch_bounds__(kThreads) void scale_row_norm_kernel(const int *__restrict__ rp,
                                                                  const double *__restrict__ val,
                                                                  const int2 *__restrict__ blk,
                                                                  double *__restrict__ out)
{
    const int b = blockIdx.x, t = threadIdx.x;
    const int2 lo = blk[b], hi = blk[b + 1];
    const int r0 = lo.x, r1 = hi.x, nr = r1 - r0;
    if (hi.y - lo.y > kLongRow * nr) {
        for (int r = r0; r < r1; ++r) {
            const int k0 = rp[r], k1 = rp[r + 1];
            double acc = 0.0;
            if (SUM) {
                if (t == 0) {
                    int k = k0;
                    for (; k + 8 <= k1; k += 8) {
                        double v[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) v[u] = fabs(val[k + u]);
#pragma unroll
                        for (int u = 0; u < 8; ++u) acc += v[u];
                    }
                    for (; k < k1; ++k) acc += fabs(val[k]);
                }
            } else {
                for (int k = k0 + t; k < k1; k += kThreads) acc = fmax(acc, fabs(val[k]));
                acc = wave_max(acc);
                __shared__ double wmax[kThreads / kWave];
                if ((t & (kWave - 1)) == 0) wmax[t / kWave] = acc;
                __syncthreads();
                if (t == 0)
                    for (int w = 1; w < kThreads / kWave; ++w) acc = fmax(acc, wmax[w]);
                __syncthreads();
            }
            if (t == 0) {
                acc = sqrt(acc);
                out[r] = acc == 0.0 ? 1.0 : acc;
            }
        }
        return;
    }
    const int i = r0 + t;
    if (i >= r1) return;
    double acc = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        double v = fabs(val[k]);
        if (SUM) acc += v;
        else if (v > acc) acc = v;
    }
    acc = sqrt(acc);
    out[i] = acc == 0.0 ? 1.0 : acc;
}

// soc_common.cuh:57-64: a cone block shares its largest factor.
__global__ void cone_block_max_kernel(int count, const int *__restrict__ off, double *__restrict__ v)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    double best = 0.0;
    for (int j = off[k]; j < off[k + 1]; ++j) best = fmax(best, v[j]);
    for (int j = off[k]; j < off[k + 1]; ++j) v[j] = best;
}

// v_k = (v_k / rdiv[row]) / cdiv[col] over the matrix's row blocks: rows
// of a stream block one per thread, the entries of a wave block's long rows
// strided over the workgroup.
__global__ __launch_bounds__(kThreads) void scale_matrix_kernel(const int *__restrict__ rp,
                                                                const int *__restrict__ ci,
                                                                double *__restrict__ val,
                                                                const int2 *__restrict__ blk,
                                                                const double *__restrict__ rdiv,
                                                                const double *__restrict__ cdiv)
{
    const int b = blockIdx.x, t = threadIdx.x;
    const int2 lo = blk[b], hi = blk[b + 1];
    const int r0 = lo.x, r1 = hi.x, nr = r1 - r0;
    if (hi.y - lo.y > kLongRow * nr) {
        for (int r = r0; r < r1; ++r) {
            const double d = rdiv[r];
            for (int k = rp[r] + t; k < rp[r + 1]; k += kThreads) {
                double v = val[k];
                v /= d;
                v /= cdiv[ci[k]];
                val[k] = v;
            }
        }
        return;
    }
    const int r = r0 + t;
    if (r >= r1) return;
    const double d = rdiv[r];
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
        double v = val[k];
        v /= d;
        v /= cdiv[ci[k]];
        val[k] = v;
    }
}

__global__ void scale_vectors_n_kernel(int n, double *__restrict__ c, double *__restrict__ l, double *__restrict__ u,
                                       double *__restrict__ col_norm, const double *__restrict__ tc)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double t = tc[j];
    col_norm[j] *= t;
    c[j] /= t;
    l[j] *= t;
    u[j] *= t;
}

__global__ void scale_vectors_m_kernel(int m, double *__restrict__ AL, double *__restrict__ AU,
                                       double *__restrict__ rhs, double *__restrict__ row_norm,
                                       const double *__restrict__ tr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double t = tr[i];
    row_norm[i] *= t;
    if (!isinf(AL[i])) AL[i] /= t;
    if (!isinf(AU[i])) AU[i] /= t;
    if (rhs) rhs[i] /= t;
}

__global__ void fill_kernel(int n, double *__restrict__ v, double value)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = value;
}

// control.c:79-114 on the device for diagonal Q. Each probe of the
// golden-section search is a sum over all n variables, so it is spread over
// the GPU, and the search looks ahead: the points a step can visit depend on
// earlier comparisons only through a branch bit, so the next kGoldenDepth
// steps have at most 2^kGoldenDepth - 1 candidate points (a binary tree,
// node (l, j) at index 2^l - 1 + j, child (l + 1, 2j + b) for branch b: b = 0
// when f1 > f2). One pass evaluates all of them -- a partial-sum kernel over
// the variables, then a one-workgroup step kernel that adds the partials in
// a fixed order, walks the tree along the branches the actual values take
// (applying the search's own updates, so its points and values are those of
// the one-probe-at-a-time search) and writes the next tree. A pass therefore
// advances up to kGoldenDepth probes; passes after convergence return at once.
constexpr int kGoldenMaxIter = 200;
constexpr int kGoldenBlocks = 256;
constexpr int kGoldenDepth = 4;
constexpr int kGoldenPoints = (1 << kGoldenDepth) - 1;
constexpr int kGoldenPasses = 2 + (kGoldenMaxIter + kGoldenDepth - 1) / kGoldenDepth;

struct GoldenState {
    double a, b, lo, hi, x1, x2, f1, f2, result;
    double pts[kGoldenPoints];
    int npts;  // points to evaluate this pass (2 for the first pass: x1, x2)
    int it, done;
};

__device__ __forceinline__ bool golden_continue(const GoldenState &g)
{
    return fabs(g.hi - g.lo) > 1e-12 * fmax(1.0, fabs(g.lo)) && g.it < kGoldenMaxIter;
}

// One search step on branch b (b = 0: f1 > f2); returns the new point. The
// value moved between f1 / f2 is moved only when the caller knows it.
__device__ __forceinline__ double golden_branch(GoldenState &g, int b, bool move_values)
{
    const double phi = (sqrt(5.0) - 1.0) / 2.0;
    if (b == 0) {
        g.lo = g.x1;
        g.x1 = g.x2;
        if (move_values) g.f1 = g.f2;
        g.x2 = g.lo + phi * (g.hi - g.lo);
        return g.x2;
    }
    g.hi = g.x2;
    g.x2 = g.x1;
    if (move_values) g.f2 = g.f1;
    g.x1 = g.hi - phi * (g.hi - g.lo);
    return g.x1;
}

// Fills g.pts with the candidate tree of the next kGoldenDepth steps from a
// state whose f1, f2 are known (or sets done when the search has ended).
__device__ void golden_plan(GoldenState &g)
{
    if (!golden_continue(g)) {
        g.done = 1;
        g.result = g.f1 < g.f2 ? g.x1 : g.x2;
        return;
    }
    GoldenState cur[1 << (kGoldenDepth - 1)], nxt[1 << (kGoldenDepth - 1)];
    bool live[1 << (kGoldenDepth - 1)], nlive[1 << (kGoldenDepth - 1)];
    // level 0: the branch is known
    cur[0] = g;
    g.pts[0] = golden_branch(cur[0], g.f1 > g.f2 ? 0 : 1, false);
    cur[0].it += 1;
    live[0] = true;
    int npts = 1;
    for (int l = 1; l < kGoldenDepth; ++l) {
        const int width = 1 << (l - 1);
        for (int j = 0; j < width; ++j) {
            for (int bb = 0; bb < 2; ++bb) {
                const int c = 2 * j + bb;
                nlive[c] = live[j] && golden_continue(cur[j]);
                nxt[c] = cur[j];
                double p = 0.0;
                if (nlive[c]) {
                    p = golden_branch(nxt[c], bb, false);
                    nxt[c].it += 1;
                }
                g.pts[(1 << l) - 1 + c] = p;  // unused when the branch cannot be reached
            }
        }
        for (int c = 0; c < 2 * width; ++c) {
            cur[c] = nxt[c];
            live[c] = nlive[c];
        }
        npts = (1 << (l + 1)) - 1;
    }
    g.npts = npts;
}

__global__ void golden_init_kernel(GoldenState *st, double a, double b)
{
    if (threadIdx.x != 0) return;
    const double phi = (sqrt(5.0) - 1.0) / 2.0;
    GoldenState g{};
    g.a = a;
    g.b = b;
    g.lo = 1e-12;
    g.hi = 1e12;
    g.x1 = g.hi - phi * (g.hi - g.lo);
    g.x2 = g.lo + phi * (g.hi - g.lo);
    g.pts[0] = g.x1;
    g.pts[1] = g.x2;
    g.npts = 2;
    *st = g;
}

__global__ __launch_bounds__(kThreads) void golden_partials_kernel(const GoldenState *__restrict__ st,
                                                                   const double *__restrict__ dq,
                                                                   const double *__restrict__ c,
                                                                   const double *__restrict__ d, int n,
                                                                   double *__restrict__ partials)
{
    if (st->done) return;
    const int np = st->npts;
    double x[kGoldenPoints], acc[kGoldenPoints];
#pragma unroll
    for (int p = 0; p < kGoldenPoints; ++p) {
        x[p] = p < np ? st->pts[p] : 0.0;
        acc[p] = 0.0;
    }
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const double qi = dq[i], ci = c[i], di = d[i];
#pragma unroll
        for (int p = 0; p < kGoldenPoints; ++p) {
            if (p < np) {
                const double t = di / (1.0 + x[p] * qi);
                acc[p] += ci * t;
            }
        }
    }
    block_reduce_store<kGoldenPoints, 0>(acc, nullptr, partials + (size_t)blockIdx.x * kGoldenPoints);
}

__global__ __launch_bounds__(kThreads) void golden_step_kernel(GoldenState *__restrict__ st,
                                                                const double *__restrict__ partials, int nblk)
{
    if (st->done) return;
    // one wavefront per point adds that point's partials (wavefront w takes
    // points w, w + 4, ...): lane-strided sums, then a fixed tree
    __shared__ double total[kGoldenPoints];
    const int np = st->npts;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (int p = wv; p < np; p += kThreads / kWave) {
        double part = 0.0;
        for (int k = lane; k < nblk; k += kWave) part += partials[(size_t)k * kGoldenPoints + p];
        part = wave_sum(part);
        if (lane == 0) total[p] = part;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    GoldenState g = *st;
    auto f_of = [&](int p) {
        const double x = g.pts[p];
        return g.a * x + g.b / x + x * x * total[p];
    };
    if (np == 2) {  // first pass: f(x1), f(x2)
        g.f1 = f_of(0);
        g.f2 = f_of(1);
    } else {
        // walk the tree along the actual branches
        int j = 0;
        for (int l = 0; l < kGoldenDepth; ++l) {
            if (!golden_continue(g)) break;
            const int bb = g.f1 > g.f2 ? 0 : 1;
            if (l > 0) j = 2 * j + bb;
            const int node = (1 << l) - 1 + j;
            (void)golden_branch(g, bb, true);
            const double f = f_of(node);
            if (bb == 0) g.f2 = f;  // the new point replaced x2
            else g.f1 = f;          // the new point replaced x1
            g.it += 1;
        }
    }
    golden_plan(g);
    *st = g;
}

// ------------------------------------------------------------ the backend

enum Vec {
    // n-sized
    V_x, V_x_bar, V_x_hat, V_w, V_w_bar, V_ATy, V_ATy_bar, V_ATdy, V_Qw, V_Qw_bar, V_Qx, V_dQw, V_QATdy, V_dx,
    V_dw, V_z_bar, V_tempv, V_diag_Q, V_fact1, V_fact2, V_fact, V_fact_M, V_last_x, V_last_Qw, V_last_w, V_last_ATy,
    V_c, V_l, V_u, V_col_norm, V_tc, V_pn0, V_pn1, V_pn2, V_carry_n,
    V_N_END,
    // m-sized
    V_y = V_N_END, V_y_bar, V_Ax, V_dy, V_s, V_last_y, V_AL, V_AU, V_row_norm, V_tr, V_soc_rhs, V_pm0, V_pm1,
    V_Rp, V_carry_m,
    V_M_END
};

static_assert(V_M_END == 50, "vector table");
constexpr int V_Rp_soc() { return V_Rp; }
const char *const kVecNames[] = {
    "x", "x_bar", "x_hat", "w", "w_bar", "ATy", "ATy_bar", "ATdy", "Qw", "Qw_bar", "Qx", "dQw", "QATdy", "dx",
    "dw", "z_bar", "tempv", "diag_Q", "fact1", "fact2", "fact", "fact_M", "last_x", "last_Qw", "last_w", "last_ATy",
    "c", "l", "u", "col_norm", "tc", "pn0", "pn1", "pn2", "carry_n",
    "y", "y_bar", "Ax", "dy", "s", "last_y", "AL", "AU", "row_norm", "tr", "soc_rhs", "pm0", "pm1", "Rp", "carry_m"};

class HipBackend final : public Backend {
public:
    ~HipBackend() override { release(); }

    HPRQPStatus setup(const HPRQPModel *model, const HPRQPParams *params);

    size_t n() const override { return n_; }
    size_t m() const override { return m_; }
    double sigma() const override { return sigma_; }
    void set_to_check(int v) override { to_check_ = v; }
    HPRQPStatus compute_residuals(HPRQPResiduals *res, size_t iter) override;
    HPRQPStatus copy_output_iterates(double *x, double *w, double *z, double *y) override;
    HPRQPStatus update_sigma(HPRQPRestart *restart, const HPRQPResiduals *res) override;
    HPRQPStatus restart_apply(HPRQPRestart *restart) override;
    double compute_m_norm() override;
    HPRQPStatus iteration_step(size_t inner) override;

    double *vec(int v) const { return v < V_N_END ? arena_ + (size_t)v * n_pad_ : arena_ + off_m_ + (size_t)(v - V_N_END) * m_pad_; }
    size_t vec_len(int v) const { return v < V_N_END ? n_ : m_; }
    const DevCsr &matrix(const char *name) const
    {
        if (!std::strcmp(name, "Q")) return Q_;
        if (!std::strcmp(name, "A")) return A_;
        return AT_;
    }
    hipStream_t stream() const { return stream_; }
    double scalar(const char *name) const;
    int gram_active() const override { return gram_ && gram_->kind() ? 1 : 0; }
    size_t gram_read(const char *name, double *out, size_t cap) const
    {
        return gram_ ? gram_->read(name, out, cap, stream_) : 0;
    }
    HPRQPResiduals last_res_{};
    size_t restart_times_ = 0;

    // Per-kernel timing of the iteration kernels with HIP events recorded on
    // this backend's own stream (profile()).
    enum Tag { T_ZXW1, T_TEMPV, T_Y, T_W2, T_ZXW1_NOC, T_QW_NOC, T_CONE_X, T_CONE_Y, T_GRAM, T_COUNT };
    HPRQPStatus profile(size_t iterations, double *mean_ms, double *bytes, size_t *launches);
    double algorithmic_bytes(int tag) const;
    static const char *tag_name(int tag)
    {
        static const char *names[] = {"zxw1_q",   "tempv_q", "y_a",    "w2_at",     "zxw1_noc_at",
                                      "qw_noc_q", "cone_x",  "cone_y", "gram_solve"};
        return names[tag];
    }

private:
    size_t n_ = 0, m_ = 0, n_pad_ = 0, m_pad_ = 0, off_m_ = 0;
    int device_ = 0;
    hipStream_t stream_ = nullptr;
    DevCsr Q_, A_, AT_;
    Panels Qp_, Ap_, ATp_;
    int *AT_perm_ = nullptr;  // A^T entry -> A entry
    int *QT_rp_ = nullptr, *QT_perm_ = nullptr;
    int *var_off_ = nullptr, *row_off_ = nullptr;
    // lanes per cone of the fused cone kernels (0: the cones do not tile
    // their range or one is longer than a wavefront; one thread per cone)
    int var_lanes_ = 0, row_lanes_ = 0;
    bool q_diag_exact_ = false;  // every row of Q holds at most its diagonal entry
    bool fused_x_off_ = std::getenv("HPRQP_CONE_LANES") && std::atoi(std::getenv("HPRQP_CONE_LANES")) == 0;
    // the cone x step runs as soc_x_diag_kernel (Q w, prefix, cones, finish)
    // the cone step's x side as one launch (soc_x_diag_kernel): Q exactly
    // diagonal, its cones (if any) tiled by lanes
    bool fused_x() const
    {
        return (num_var_cones_ || num_row_cones_ || gram_) && (!num_var_cones_ || var_lanes_) && q_diag_exact_ &&
               Qp_.P <= 1 && !fused_x_off_;
    }
    template <class Fin>
    void cone_step(int lanes, size_t count, const int *off, const double *in, double *out, const Fin &fin);
    size_t num_var_cones_ = 0, num_row_cones_ = 0;
    int gram_mode_ = 0;                 // 2: the equality-row metric is requested (AUTO resolved)
    std::unique_ptr<GramMetric> gram_;  // set when the metric was accepted
    double *arena_ = nullptr;
    double *partials_ = nullptr;  // [4 regions][max_grid * kMaxTerms]
    size_t region_ = 0;
    double *red_ = nullptr;       // device reduction results
    struct GoldenState *golden_ = nullptr;  // device state of the diagonal sigma search
    bool aty_bar_valid_ = false;  // ATy_bar holds A^T y_bar of the current y_bar
    double *golden_part_ = nullptr;          // its per-workgroup partial sums
    hipGraphExec_t golden_graph_ = nullptr;  // its captured pass chain
    double *red_host_ = nullptr;  // pinned mirror
    std::vector<double> col_norm_h_, row_norm_h_;
    double sigma_ = 1.0, lambda_A_ = 0.0, lambda_Q_ = 0.0;
    int power_steps_A_ = 0, power_steps_Q_ = 0;  // products taken by the power iterations
    int Q_is_diag_ = 0, noC_ = 0, to_check_ = 1;
    double norm_b_org_ = 0.0, norm_c_org_ = 0.0, obj_constant_ = 0.0;
    double b_scale_ = 1.0, c_scale_ = 1.0;
    bool profiling_ = false;
    std::vector<hipEvent_t> ev_pool_;
    std::vector<int> ev_tags_;
    size_t ev_used_ = 0;
    void prof_begin(int tag)
    {
        if (ev_used_ + 2 > ev_pool_.size()) return;
        (void)hipEventRecord(ev_pool_[ev_used_], stream_);
        ev_tags_[ev_used_] = tag;
        ++ev_used_;
    }
    void prof_end()
    {
        if (ev_used_ == 0 || ev_used_ + 1 > ev_pool_.size() || ev_tags_[ev_used_ - 1] < 0) return;
        (void)hipEventRecord(ev_pool_[ev_used_], stream_);
        ev_tags_[ev_used_] = -1;
        ++ev_used_;
    }

    void release();
    HPRQPStatus upload(const HPRQPModel *model);
    HPRQPStatus scale(const HPRQPParams *params);
    HPRQPStatus workspace_setup(const HPRQPModel *model);
    HPRQPStatus power_iteration(bool q_matrix, double *lambda, int *steps);
    HPRQPStatus soc_iteration_step(size_t inner);
    HPRQPStatus soc_residuals(HPRQPResiduals *res);
    size_t lin_n_ = 0, lin_m_ = 0;

    double *region(int r) const { return partials_ + (size_t)r * region_; }

    const Panels &panels_of(const DevCsr &M) const { return &M == &Q_ ? Qp_ : (&M == &A_ ? Ap_ : ATp_); }
    // nblocks of the launch that carries the reduction terms (the last panel)
    int red_blocks(const DevCsr &M) const
    {
        const Panels &pn = panels_of(M);
        return pn.P > 1 ? pn.part[pn.P - 1].nblocks : M.nblocks;
    }
    template <class Epi>
    HPRQPStatus spmv(const DevCsr &M, const double *x, Epi epi, int reg = 0, int tag = -1)
    {
        if (M.nrow == 0) return HPRQP_STATUS_OK;
        if (Epi::NS + Epi::NM > 0) {
            if (pending_regions_ >> reg & 1u) flush_red();  // its partial rows are still queued
            epi.partials = region(reg);
        }
        const bool timed = profiling_ && tag >= 0;
        if (timed) prof_begin(tag);
        const Panels &pn = panels_of(M);
        const size_t pad = pn.lds_pad;
        if (pn.P <= 1) {
            spmv_rows_kernel<Epi><<<M.nblocks, kThreads, pad, stream_>>>(M.rp, M.ci, M.val, x, M.blocks, epi);
        } else {
            double *carry = M.nrow == (int)n_ ? vec(V_carry_n) : vec(V_carry_m);
            for (int p = 0; p + 1 < pn.P; ++p) {
                const DevCsr &d = pn.part[p];
                if (p == 0)
                    spmv_rows_kernel<EpiCarry, false><<<d.nblocks, kThreads, pad, stream_>>>(
                        d.rp, d.ci, d.val, x, d.blocks, EpiCarry{{}, carry}, carry);
                else
                    spmv_rows_kernel<EpiCarry, true><<<d.nblocks, kThreads, pad, stream_>>>(
                        d.rp, d.ci, d.val, x, d.blocks, EpiCarry{{}, carry}, carry);
            }
            const DevCsr &d = pn.part[pn.P - 1];
            spmv_rows_kernel<Epi, true><<<d.nblocks, kThreads, pad, stream_>>>(d.rp, d.ci, d.val, x, d.blocks, epi,
                                                                              carry);
        }
        if (timed) prof_end();
        return HPRQP_STATUS_OK;
    }
    HPRQPStatus autotune_panels();
    template <class F>
    HPRQPStatus map(size_t len, F f, int reg = 0)
    {
        if (len == 0) return HPRQP_STATUS_OK;
        if (F::NS + F::NM > 0) {
            if (pending_regions_ >> reg & 1u) flush_red();  // its partial rows are still queued
            f.partials = region(reg);
        }
        map_kernel<F><<<grid_for(len), kThreads, 0, stream_>>>((int)len, f);
        return HPRQP_STATUS_OK;
    }
    // Reduce region `reg` written by a launch of `nblk` workgroups with
    // ns sums and nm maxima into red_[dst .. dst+ns+nm).
    void finish(int reg, int nblk, int ns, int nm, int dst)
    {
        if (nblk <= 0 || ns + nm == 0) return;
        reduce_partials_kernel<<<ns + nm, kThreads, 0, stream_>>>(region(reg), nblk, ns + nm, ns, red_ + dst);
    }
    // A check's reductions are finished together: finish_later() queues a
    // region (its partial rows must stay untouched until the flush) and
    // flush_red() finishes the queue in one launch; read_red() flushes first.
    RedJobs pending_{};
    unsigned pending_regions_ = 0;  // regions whose partial rows a queued job reads
    void finish_later(int reg, int nblk, int ns, int nm, int dst)
    {
        if (nblk <= 0 || ns + nm == 0) return;
        if (pending_.count == kMaxRedJobs) flush_red();
        pending_regions_ |= 1u << reg;
        RedJobs &q = pending_;
        const int j = q.count++;
        q.partials[j] = region(reg);
        q.out[j] = red_ + dst;
        q.nblk[j] = nblk;
        q.nterms[j] = ns + nm;
        q.ns[j] = ns;
        q.first[j] = j ? q.first[j - 1] + q.nterms[j - 1] : 0;
    }
    void flush_red()
    {
        if (pending_.count == 0) return;
        const unsigned blocks = (unsigned)(pending_.first[pending_.count - 1] + pending_.nterms[pending_.count - 1]);
        reduce_jobs_kernel<<<blocks, kThreads, 0, stream_>>>(pending_);
        pending_.count = 0;
        pending_regions_ = 0;
    }
    HPRQPStatus read_red(int count)
    {
        flush_red();
        HPRQP_HIP(hipMemcpyAsync(red_host_, red_, sizeof(double) * count, hipMemcpyDeviceToHost, stream_));
        HPRQP_HIP(hipStreamSynchronize(stream_));
        return HPRQP_STATUS_OK;
    }
    void zero_red(int from, int count)
    {
        fill_kernel<<<1, kThreads, 0, stream_>>>(count, red_ + from, 0.0);
    }
    int n_grid() const { return grid_for(n_); }
    int m_grid() const { return grid_for(m_); }
};

void HipBackend::release()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    gram_.reset();
    free_panels(Qp_);
    free_panels(Ap_);
    free_panels(ATp_);
    free_csr(Q_);
    free_csr(A_);
    free_csr(AT_);
    if (AT_perm_) (void)hipFree(AT_perm_);
    if (QT_rp_) (void)hipFree(QT_rp_);
    if (QT_perm_) (void)hipFree(QT_perm_);
    if (var_off_) (void)hipFree(var_off_);
    if (row_off_) (void)hipFree(row_off_);
    if (arena_) (void)hipFree(arena_);
    if (partials_) (void)hipFree(partials_);
    if (red_) (void)hipFree(red_);
    if (golden_graph_) (void)hipGraphExecDestroy(golden_graph_);
    golden_graph_ = nullptr;
    if (golden_) (void)hipFree(golden_);
    golden_ = nullptr;
    if (golden_part_) (void)hipFree(golden_part_);
    golden_part_ = nullptr;
    if (red_host_) (void)hipHostFree(red_host_);
    if (stream_) (void)hipStreamDestroy(stream_);
    AT_perm_ = QT_rp_ = QT_perm_ = var_off_ = row_off_ = nullptr;
    arena_ = partials_ = red_ = red_host_ = nullptr;
    stream_ = nullptr;
}

HPRQPStatus upload_csr(DevCsr &d, const HPRQPMatrixCSR &h, hipStream_t s)
{
    if (h.nnz > (size_t)INT32_MAX || h.nrow >= (size_t)INT32_MAX || h.ncol >= (size_t)INT32_MAX)
        return HPRQP_STATUS_INVALID_ARGUMENT;
    d.nrow = (int)h.nrow;
    d.ncol = (int)h.ncol;
    d.nnz = (int)h.nnz;
    std::vector<int> rp(h.nrow + 1), ci(h.nnz);
    for (size_t r = 0; r <= h.nrow; ++r) rp[r] = (int)h.row_ptr[r];
    for (size_t k = 0; k < h.nnz; ++k) {
        if (h.col_ind[k] >= h.ncol) return HPRQP_STATUS_INVALID_ARGUMENT;
        ci[k] = (int)h.col_ind[k];
    }
    HPRQP_HIP(hipMalloc(&d.rp, sizeof(int) * (h.nrow + 1)));
    HPRQP_HIP(hipMalloc(&d.ci, sizeof(int) * (h.nnz ? h.nnz : 1)));
    HPRQP_HIP(hipMalloc(&d.val, sizeof(double) * (h.nnz ? h.nnz : 1)));
    HPRQP_HIP(hipMemcpyAsync(d.rp, rp.data(), sizeof(int) * (h.nrow + 1), hipMemcpyHostToDevice, s));
    if (h.nnz) {
        HPRQP_HIP(hipMemcpyAsync(d.ci, ci.data(), sizeof(int) * h.nnz, hipMemcpyHostToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(d.val, h.values, sizeof(double) * h.nnz, hipMemcpyHostToDevice, s));
    }
    HPRQP_HIP(hipStreamSynchronize(s));
    return HPRQP_STATUS_OK;
}

HPRQPStatus HipBackend::upload(const HPRQPModel *md)
{
    n_ = md->Q.ncol;
    m_ = md->A.nrow;
    n_pad_ = (n_ + 31) & ~(size_t)31;
    m_pad_ = (m_ + 31) & ~(size_t)31;
    if (n_pad_ == 0) n_pad_ = 32;
    if (m_pad_ == 0) m_pad_ = 32;
    off_m_ = n_pad_ * V_N_END;
    const size_t total = off_m_ + m_pad_ * (V_M_END - V_N_END);
    HPRQP_HIP(hipMalloc(&arena_, sizeof(double) * total));
    fill_kernel<<<grid_for(total), kThreads, 0, stream_>>>((int)std::min(total, (size_t)INT32_MAX), arena_, 0.0);
    HPRQPStatus st = upload_csr(Q_, md->Q, stream_);
    if (st == HPRQP_STATUS_OK) st = upload_csr(A_, md->A, stream_);
    if (st != HPRQP_STATUS_OK) return st;
    HPRQP_HIP(hipMemcpyAsync(vec(V_c), md->c, sizeof(double) * n_, hipMemcpyHostToDevice, stream_));
    HPRQP_HIP(hipMemcpyAsync(vec(V_l), md->l, sizeof(double) * n_, hipMemcpyHostToDevice, stream_));
    HPRQP_HIP(hipMemcpyAsync(vec(V_u), md->u, sizeof(double) * n_, hipMemcpyHostToDevice, stream_));
    if (m_) {
        HPRQP_HIP(hipMemcpyAsync(vec(V_AL), md->AL, sizeof(double) * m_, hipMemcpyHostToDevice, stream_));
        HPRQP_HIP(hipMemcpyAsync(vec(V_AU), md->AU, sizeof(double) * m_, hipMemcpyHostToDevice, stream_));
        if (md->num_row_cones)
            HPRQP_HIP(hipMemcpyAsync(vec(V_soc_rhs), md->soc_rhs, sizeof(double) * m_, hipMemcpyHostToDevice, stream_));
    }
    num_var_cones_ = md->num_var_cones;
    num_row_cones_ = md->num_row_cones;
    lin_n_ = num_var_cones_ ? md->var_cone_offsets[0] : n_;
    lin_m_ = num_row_cones_ ? md->row_cone_offsets[0] : m_;
    // lanes per cone for the fused kernels: cones that tile [lin, len) in
    // order, each at most a wavefront long
    auto lanes_for = [](const size_t *o, size_t count, size_t len) {
        size_t longest = 0;
        if (o[count] != len) return 0;
        for (size_t k = 0; k < count; ++k) {
            if (o[k + 1] < o[k]) return 0;
            longest = std::max(longest, o[k + 1] - o[k]);
        }
        if (longest > (size_t)kWave) return 0;
        int l = 1;
        while ((size_t)l < longest) l <<= 1;
        return l;
    };
    var_lanes_ = num_var_cones_ ? lanes_for(md->var_cone_offsets, num_var_cones_, n_) : 0;
    q_diag_exact_ = true;
    for (size_t r = 0; r < md->Q.nrow && q_diag_exact_; ++r) {
        const size_t e0 = md->Q.row_ptr[r], e1 = md->Q.row_ptr[r + 1];
        if (e1 - e0 > 1 || (e1 > e0 && md->Q.col_ind[e0] != r)) q_diag_exact_ = false;
    }
    row_lanes_ = num_row_cones_ ? lanes_for(md->row_cone_offsets, num_row_cones_, m_) : 0;
    if (const char *v = std::getenv("HPRQP_CONE_LANES"); v && std::atoi(v) == 0) var_lanes_ = row_lanes_ = 0;
    if (num_var_cones_) {
        std::vector<int> off(num_var_cones_ + 1);
        for (size_t k = 0; k <= num_var_cones_; ++k) off[k] = (int)md->var_cone_offsets[k];
        HPRQP_HIP(hipMalloc(&var_off_, sizeof(int) * off.size()));
        HPRQP_HIP(hipMemcpyAsync(var_off_, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, stream_));
    }
    if (num_row_cones_) {
        std::vector<int> off(num_row_cones_ + 1);
        for (size_t k = 0; k <= num_row_cones_; ++k) off[k] = (int)md->row_cone_offsets[k];
        HPRQP_HIP(hipMalloc(&row_off_, sizeof(int) * off.size()));
        HPRQP_HIP(hipMemcpyAsync(row_off_, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, stream_));
    }
    // A^T on the device, with the permutation kept for re-gathering values.
    AT_.nrow = (int)n_;
    AT_.ncol = (int)m_;
    AT_.nnz = A_.nnz;
    HPRQP_HIP(hipMalloc(&AT_.rp, sizeof(int) * (n_ + 1)));
    HPRQP_HIP(hipMalloc(&AT_.ci, sizeof(int) * (A_.nnz ? A_.nnz : 1)));
    HPRQP_HIP(hipMalloc(&AT_.val, sizeof(double) * (A_.nnz ? A_.nnz : 1)));
    HPRQP_HIP(hipMalloc(&AT_perm_, sizeof(int) * (A_.nnz ? A_.nnz : 1)));
    st = device_transpose(A_.nrow, (int)n_, A_.nnz, A_.rp, A_.ci, A_.val, AT_.rp, AT_.ci, AT_.val, AT_perm_, stream_);
    if (st != HPRQP_STATUS_OK) return st;
    // Q^T permutation (structure of Q^T equals Q's only when Q is structurally
    // symmetric; keep its own row pointer to be general).
    HPRQP_HIP(hipMalloc(&QT_rp_, sizeof(int) * (n_ + 1)));
    HPRQP_HIP(hipMalloc(&QT_perm_, sizeof(int) * (Q_.nnz ? Q_.nnz : 1)));
    st = device_transpose(Q_.nrow, (int)n_, Q_.nnz, Q_.rp, Q_.ci, Q_.val, QT_rp_, nullptr, nullptr, QT_perm_, stream_);
    if (st != HPRQP_STATUS_OK) return st;
    st = upload_blocks(Q_, stream_);
    if (st == HPRQP_STATUS_OK) st = upload_blocks(A_, stream_);
    if (st == HPRQP_STATUS_OK) st = upload_blocks(AT_, stream_);
    if (st != HPRQP_STATUS_OK) return st;
    size_t max_grid = std::max({(size_t)Q_.nblocks, (size_t)A_.nblocks, (size_t)AT_.nblocks, (size_t)n_grid(),
                                (size_t)m_grid(), (size_t)1});
    region_ = max_grid * kMaxTerms;
    HPRQP_HIP(hipMalloc(&partials_, sizeof(double) * region_ * 4));
    HPRQP_HIP(hipMalloc(&red_, sizeof(double) * kRedLen));
    HPRQP_HIP(hipHostMalloc(reinterpret_cast<void **>(&red_host_), sizeof(double) * kRedLen, hipHostMallocDefault));
    return HPRQP_STATUS_OK;
}

// scale.c:231-318, plus the SOC block uniformisation of soc_common.cuh:57-75.
HPRQPStatus HipBackend::scale(const HPRQPParams *params)
{
    const int n = (int)n_, m = (int)m_;
    fill_kernel<<<grid_for(n_), kThreads, 0, stream_>>>(n, vec(V_col_norm), 1.0);
    if (m) fill_kernel<<<grid_for(m_), kThreads, 0, stream_>>>(m, vec(V_row_norm), 1.0);
    double *rhs = num_row_cones_ ? vec(V_soc_rhs) : nullptr;
    if (params->use_scaling) {
        for (int pass = 0; pass < 11; ++pass) {
            const bool sum = pass == 10;
            if (sum)
                scale_col_norm_kernel<true><<<grid_for(n_), kThreads, 0, stream_>>>(
                    n, AT_.rp, AT_perm_, A_.val, QT_rp_, QT_perm_, Q_.val, vec(V_tc));
            else
                scale_col_norm_kernel<false><<<grid_for(n_), kThreads, 0, stream_>>>(
                    n, AT_.rp, AT_perm_, A_.val, QT_rp_, QT_perm_, Q_.val, vec(V_tc));
            if (m) {
                if (sum)
                    scale_row_norm_kernel<true><<<A_.nblocks, kThreads, 0, stream_>>>(A_.rp, A_.val, A_.blocks,
                                                                                     vec(V_tr));
                else
                    scale_row_norm_kernel<false><<<A_.nblocks, kThreads, 0, stream_>>>(A_.rp, A_.val, A_.blocks,
                                                                                      vec(V_tr));
            }
            if (num_var_cones_)
                cone_block_max_kernel<<<grid_for(num_var_cones_), kThreads, 0, stream_>>>((int)num_var_cones_,
                                                                                         var_off_, vec(V_tc));
            if (num_row_cones_)
                cone_block_max_kernel<<<grid_for(num_row_cones_), kThreads, 0, stream_>>>((int)num_row_cones_,
                                                                                         row_off_, vec(V_tr));
            if (Q_.nrow)
                scale_matrix_kernel<<<Q_.nblocks, kThreads, 0, stream_>>>(Q_.rp, Q_.ci, Q_.val, Q_.blocks, vec(V_tc),
                                                                          vec(V_tc));
            if (m)
                scale_matrix_kernel<<<A_.nblocks, kThreads, 0, stream_>>>(A_.rp, A_.ci, A_.val, A_.blocks, vec(V_tr),
                                                                          vec(V_tc));
            scale_vectors_n_kernel<<<grid_for(n_), kThreads, 0, stream_>>>(n, vec(V_c), vec(V_l), vec(V_u),
                                                                           vec(V_col_norm), vec(V_tc));
            if (m)
                scale_vectors_m_kernel<<<grid_for(m_), kThreads, 0, stream_>>>(m, vec(V_AL), vec(V_AU), rhs,
                                                                               vec(V_row_norm), vec(V_tr));
        }
    }
    if (A_.nnz)
        gather_values_kernel<<<grid_for(A_.nnz), kThreads, 0, stream_>>>(A_.nnz, AT_perm_, A_.val, AT_.val);
    HPRQP_HIP(hipGetLastError());
    col_norm_h_.resize(n_);
    row_norm_h_.resize(m_);
    HPRQP_HIP(hipMemcpyAsync(col_norm_h_.data(), vec(V_col_norm), sizeof(double) * n_, hipMemcpyDeviceToHost, stream_));
    if (m_)
        HPRQP_HIP(hipMemcpyAsync(row_norm_h_.data(), vec(V_row_norm), sizeof(double) * m_, hipMemcpyDeviceToHost, stream_));
    HPRQP_HIP(hipStreamSynchronize(stream_));
    return HPRQP_STATUS_OK;
}

// Panel-count autotuner: for Q, A and A^T try 1, 2 and 4 column panels
// (only when the gathered vector exceeds 1 MiB and panels stay >= 32k
// columns wide), time five plain products of each with HIP events on this
// stream, keep the fastest, more panels only with a 3 % margin.
// HPRQP_PANELS=1|2|4 forces a count.
HPRQPStatus HipBackend::autotune_panels()
{
    const char *force = std::getenv("HPRQP_PANELS");
    const int forced = force ? std::atoi(force) : 0;
    // Residency cap for the products (see below): a dynamic LDS request of
    // a third of the CU's LDS minus the kernel's static LDS (the 16 KiB
    // product stage and a few reduction slots) holds three workgroups per
    // CU. HPRQP_OCCUPANCY_CAP=0 never applies it, =1 always where eligible.
    const char *cap_env = std::getenv("HPRQP_OCCUPANCY_CAP");
    const int cap_forced = cap_env ? std::atoi(cap_env) : -1;
    int lds_cu = 0, lds_block = 0, cus = 0;  // a failed query leaves the cap off
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device_) != hipSuccess ||
        hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device_) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_) != hipSuccess) {
        lds_cu = lds_block = cus = 0;
        (void)hipGetLastError();  // not an error of this solve: clear it
    }
    constexpr int kStaticLds = 17 * 1024;  // above every spmv_rows_kernel's static LDS
    int pad3 = (lds_cu / 3 - kStaticLds) & ~255;
    if (cus <= 0 || pad3 <= 0 || kStaticLds + pad3 > lds_block || lds_cu / (kStaticLds + pad3) != 3 ||
        lds_cu / (kStaticLds - 1024 + pad3) != 3)
        pad3 = 0;
    struct Item {
        const DevCsr *M;
        Panels *pn;
        double *x, *y;
    } items[3] = {{&Q_, &Qp_, vec(V_pn0), vec(V_pn1)},
                  {&A_, &Ap_, vec(V_pn0), vec(V_pm0)},
                  {&AT_, &ATp_, vec(V_pm1), vec(V_pn2)}};
    EventPair ev;
    HPRQP_HIP(ev.create());
    hipEvent_t e0 = ev.a, e1 = ev.b;
    for (Item &it : items) {
        const DevCsr &M = *it.M;
        std::vector<int> cands = {1};
        if (forced == 1 || forced == 2 || forced == 4) {
            cands = {forced};
        } else if (&M == &Q_ && q_diag_exact_) {
            // at most the diagonal entry per row: the gather is a coalesced
            // stream that no panel can shorten (and the cone models' fused
            // x step wants one panel)
        } else if ((size_t)M.ncol * sizeof(double) > ((size_t)1 << 20) && M.nnz > 0) {
            if (M.ncol >= 2 * 32768) cands.push_back(2);
            if (M.ncol >= 4 * 32768) cands.push_back(4);
        }
        auto probe_ms = [&](float *out) -> HPRQPStatus {
            float ms_min = 1e30f;
            for (int rep = 0; rep < 5; ++rep) {
                HPRQP_HIP(hipEventRecord(e0, stream_));
                spmv(M, it.x, EpiStore{{}, it.y});
                HPRQP_HIP(hipEventRecord(e1, stream_));
                HPRQP_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                HPRQP_HIP(hipEventElapsedTime(&ms, e0, e1));
                ms_min = std::min(ms_min, ms);
            }
            *out = ms_min;
            return HPRQP_STATUS_OK;
        };
        int best = 1;
        float best_ms = 1e30f;
        for (int P : cands) {
            HPRQPStatus st = build_panels(M, P, *it.pn, stream_);
            if (st == HPRQP_STATUS_OK) st = fill_panel_values(M, *it.pn, stream_);
            if (st != HPRQP_STATUS_OK) return st;
            if (cands.size() == 1) {
                best = P;
                break;
            }
            float ms_min = 1e30f;
            st = probe_ms(&ms_min);
            if (st != HPRQP_STATUS_OK) return st;
            // more panels must win by 3 %: the loop's products carry heavier
            // epilogues than this probe, and every panel adds a launch and a
            // carry pass
            if (P > 1) ms_min *= 1.03f;
            if (ms_min < best_ms) {
                best_ms = ms_min;
                best = P;
            }
        }
        if (it.pn->P != best) {
            HPRQPStatus st = build_panels(M, best, *it.pn, stream_);
            if (st == HPRQP_STATUS_OK) st = fill_panel_values(M, *it.pn, stream_);
            if (st != HPRQP_STATUS_OK) return st;
        }
        // Occupancy: the products are bound by the L2 request rate of their
        // random gathers; past about three resident workgroups per CU the
        // extra ones mostly evict each other's lines of x. For a matrix with
        // many more row blocks than the capped residency holds, the cap is
        // kept when it beats the uncapped product by 1 % on this probe (it
        // changes no result: the same blocks run in the same order).
        it.pn->lds_pad = 0;
        int launch_blocks = M.nblocks;
        if (it.pn->P > 1) {
            launch_blocks = 0;
            for (int p = 0; p < it.pn->P; ++p) launch_blocks = std::max(launch_blocks, it.pn->part[p].nblocks);
        }
        if (pad3 > 0 && cap_forced != 0 && M.nnz > 0 && launch_blocks >= 6 * cus) {
            float plain = 0.0f, capped = 0.0f;
            HPRQPStatus st = cap_forced == 1 ? HPRQP_STATUS_OK : probe_ms(&plain);
            if (st != HPRQP_STATUS_OK) return st;
            it.pn->lds_pad = (unsigned)pad3;
            if (cap_forced != 1) {
                st = probe_ms(&capped);
                if (st != HPRQP_STATUS_OK) return st;
                if (!(capped < 0.99f * plain)) it.pn->lds_pad = 0;
            }
        }
    }
    // the reduction partials must cover the largest panel launch
    size_t need = region_ / kMaxTerms;
    for (const Panels *pn : {&Qp_, &Ap_, &ATp_})
        for (int p = 0; p < pn->P; ++p) need = std::max(need, (size_t)pn->part[p].nblocks);
    if (need * kMaxTerms > region_) {
        HPRQP_HIP(hipStreamSynchronize(stream_));
        HPRQP_HIP(hipFree(partials_));
        region_ = need * kMaxTerms;
        HPRQP_HIP(hipMalloc(&partials_, sizeof(double) * region_ * 4));
    }
    HPRQP_HIP(hipStreamSynchronize(stream_));
    return HPRQP_STATUS_OK;
}

// eigs.c:19-153: power iteration on A A^T (q_matrix = false) or Q.
//
// The steps are issued in runs of kPowRun without waiting for the host: the
// residual pass of each step also forms the next step's q = z / ||z|| from
// ||z||^2 on the device (the same correctly rounded square root the host
// takes), and every step leaves (lambda, ||z||^2, ||z - lambda q||^2) in its
// own slot of red_. After a run the host reads the slots and replays the
// stopping rule step by step, so lambda and the step count are those of the
// one-step-at-a-time loop; the steps issued past the stopping step only
// overwrite the scratch vectors.
HPRQPStatus HipBackend::power_iteration(bool q_matrix, double *lambda, int *steps)
{
    const size_t len = q_matrix ? n_ : m_;
    *steps = 0;
    if (len == 0) {
        *lambda = 0.0;
        return HPRQP_STATUS_OK;
    }
    double *z = q_matrix ? vec(V_pn0) : vec(V_pm0);
    double *q = q_matrix ? vec(V_pn1) : vec(V_pm1);
    double *tmp = vec(V_pn2);
    const int g = grid_for(len);
    map(len, MapProbe{nullptr, z, 0.0, (double)len + 1.0}, 0);
    finish(0, g, 1, 0, 0);
    HPRQPStatus st = read_red(1);
    if (st != HPRQP_STATUS_OK) return st;
    double qn = std::sqrt(red_host_[0]);
    double lam = 1.0;
    const DevCsr &M = q_matrix ? Q_ : A_;
    map(len, MapScaleDivNorm{{}, q, z, red_});  // the first q, from the probe's ||z||^2
    constexpr int kMaxSteps = 5000;
    bool done = false;
    for (int it0 = 0; it0 < kMaxSteps && !done; it0 += kPowRun) {
        if (qn <= 1e-16) {
            lam = 0.0;
            break;
        }
        const int run = std::min(kPowRun, kMaxSteps - it0);
        for (int j = 0; j < run; ++j) {
            const int slot = kPowSlots + 3 * j;
            if (!q_matrix) spmv(AT_, q, EpiStore{{}, tmp});
            spmv(M, q_matrix ? q : tmp, EpiPowerZ{nullptr, z, q}, 0);
            finish(0, red_blocks(M), 2, 0, slot);
            map(len, MapPowerResid{nullptr, q, z, red_ + slot}, 1);
            finish(1, g, 1, 0, slot + 2);
        }
        HPRQP_HIP(hipGetLastError());
        HPRQP_HIP(hipMemcpyAsync(red_host_ + kPowSlots, red_ + kPowSlots, sizeof(double) * 3 * run,
                                 hipMemcpyDeviceToHost, stream_));
        HPRQP_HIP(hipStreamSynchronize(stream_));
        for (int j = 0; j < run; ++j) {
            if (qn <= 1e-16) {  // the step after this one divided by ~0: its values are discarded
                lam = 0.0;
                done = true;
                break;
            }
            const double *r = red_host_ + kPowSlots + 3 * j;
            *steps = it0 + j + 1;
            lam = r[0];
            const double zn = std::sqrt(r[1]);
            const double err = std::sqrt(r[2]) / (zn + std::fabs(lam) + 1e-16);
            if (err < 1e-4) {
                done = true;
                break;
            }
            qn = zn;
        }
    }
    *lambda = lam < 0.0 ? 0.0 : lam;
    return HPRQP_STATUS_OK;
}

template <class Fin>
void HipBackend::cone_step(int lanes, size_t count, const int *off, const double *in, double *out, const Fin &fin)
{
    const unsigned grid = (unsigned)((count * (size_t)lanes + kThreads - 1) / kThreads);
    switch (lanes) {
    case 1: soc_cone_group_kernel<1, Fin><<<grid, kThreads, 0, stream_>>>((int)count, off, in, out, fin); break;
    case 2: soc_cone_group_kernel<2, Fin><<<grid, kThreads, 0, stream_>>>((int)count, off, in, out, fin); break;
    case 4: soc_cone_group_kernel<4, Fin><<<grid, kThreads, 0, stream_>>>((int)count, off, in, out, fin); break;
    case 8: soc_cone_group_kernel<8, Fin><<<grid, kThreads, 0, stream_>>>((int)count, off, in, out, fin); break;
    case 16: soc_cone_group_kernel<16, Fin><<<grid, kThreads, 0, stream_>>>((int)count, off, in, out, fin); break;
    case 32: soc_cone_group_kernel<32, Fin><<<grid, kThreads, 0, stream_>>>((int)count, off, in, out, fin); break;
    default: soc_cone_group_kernel<64, Fin><<<grid, kThreads, 0, stream_>>>((int)count, off, in, out, fin); break;
    }
}

// workspace.c:191-248 (+ control.cu:189-220 for cone models)
HPRQPStatus HipBackend::workspace_setup(const HPRQPModel *md)
{
    // sigma0 from the scaled c and bounds, summed on the host in index order.
    std::vector<double> c(n_), al(m_), au(m_), rhs;
    HPRQP_HIP(hipMemcpyAsync(c.data(), vec(V_c), sizeof(double) * n_, hipMemcpyDeviceToHost, stream_));
    if (m_) {
        HPRQP_HIP(hipMemcpyAsync(al.data(), vec(V_AL), sizeof(double) * m_, hipMemcpyDeviceToHost, stream_));
        HPRQP_HIP(hipMemcpyAsync(au.data(), vec(V_AU), sizeof(double) * m_, hipMemcpyDeviceToHost, stream_));
        if (num_row_cones_) {
            rhs.resize(m_);
            HPRQP_HIP(hipMemcpyAsync(rhs.data(), vec(V_soc_rhs), sizeof(double) * m_, hipMemcpyDeviceToHost, stream_));
        }
    }
    HPRQP_HIP(hipStreamSynchronize(stream_));
    double nc2 = 0.0, nb2 = 0.0;
    for (size_t i = 0; i < n_; ++i) nc2 += c[i] * c[i];
    for (size_t i = 0; i < m_; ++i) {
        double a = std::isinf(al[i]) ? 0.0 : std::fabs(al[i]);
        double b = std::isinf(au[i]) ? 0.0 : std::fabs(au[i]);
        double v = a > b ? a : b;
        nb2 += v * v;
    }
    double nc = std::sqrt(nc2), nb = std::sqrt(nb2);
    if (num_row_cones_) {
        double r2 = 0.0;
        for (size_t i = 0; i < m_; ++i) r2 += rhs[i] * rhs[i];
        nb = std::fmax(nb, std::sqrt(r2));
    }
    sigma_ = (nb > 1e-16 && nc > 1e-16 && nb < 1e16 && nc < 1e16) ? nb / nc : 1.0;
    size_t free_count = 0;
    for (size_t i = 0; i < n_; ++i)
        if (std::isinf(md->l[i]) && std::isinf(md->u[i])) ++free_count;
    noC_ = n_ > 0 && 5 * free_count > 4 * n_;
    // diag(Q) and the diagonal test (off-diagonal magnitudes <= 1e-15)
    map(n_, DiagInfo{nullptr, vec(V_diag_Q), Q_.rp, Q_.ci, Q_.val}, 0);
    finish(0, n_grid(), 0, 1, 0);
    HPRQPStatus st = read_red(1);
    if (st != HPRQP_STATUS_OK) return st;
    Q_is_diag_ = n_ == 0 || red_host_[0] <= 1e-15;
    if (Q_is_diag_) map(n_, MapDiagFactors{{}, vec(V_fact1), vec(V_fact2), vec(V_fact), vec(V_fact_M), vec(V_diag_Q), sigma_});
    const bool cones = num_var_cones_ || num_row_cones_;
    SetupClock clock(stream_);
    st = power_iteration(false, &lambda_A_, &power_steps_A_);
    if (st != HPRQP_STATUS_OK) return st;
    clock.mark("  power iteration on A A'");
    lambda_A_ *= 1.05;
    if (Q_is_diag_) {
        std::vector<double> d(n_);
        HPRQP_HIP(hipMemcpyAsync(d.data(), vec(V_diag_Q), sizeof(double) * n_, hipMemcpyDeviceToHost, stream_));
        HPRQP_HIP(hipStreamSynchronize(stream_));
        lambda_Q_ = 0.0;
        for (size_t i = 0; i < n_; ++i) {
            double v = cones ? std::fabs(d[i]) : d[i];
            if (v > lambda_Q_) lambda_Q_ = v;
        }
    } else {
        st = power_iteration(true, &lambda_Q_, &power_steps_Q_);
        if (st != HPRQP_STATUS_OK) return st;
        clock.mark("  power iteration on Q");
        lambda_Q_ *= 1.05;
    }
    if (cones) {
        noC_ = 0;
        if (m_ > 0 && lambda_A_ <= 0.0) lambda_A_ = 1.0;
    }
    if (gram_mode_ == 2) {
        // control.cu:222-239: the metric is set up on the scaled model; a
        // declined metric leaves sigma, lambda_A and noC as they are
        std::vector<int> rp(m_ + 1), ci(A_.nnz);
        std::vector<double> val(A_.nnz);
        HPRQP_HIP(hipMemcpyAsync(rp.data(), A_.rp, sizeof(int) * (m_ + 1), hipMemcpyDeviceToHost, stream_));
        if (A_.nnz) {
            HPRQP_HIP(hipMemcpyAsync(ci.data(), A_.ci, sizeof(int) * A_.nnz, hipMemcpyDeviceToHost, stream_));
            HPRQP_HIP(hipMemcpyAsync(val.data(), A_.val, sizeof(double) * A_.nnz, hipMemcpyDeviceToHost, stream_));
        }
        HPRQP_HIP(hipStreamSynchronize(stream_));
        GramInputs in;
        in.n = (int)n_;
        in.m = (int)m_;
        in.rp = rp.data();
        in.ci = ci.data();
        in.val = val.data();
        in.AL = al.data();
        in.AU = au.data();
        in.c = c.data();
        in.soc_rhs = num_row_cones_ ? rhs.data() : nullptr;
        in.num_var_cones = num_var_cones_;
        in.num_row_cones = num_row_cones_;
        in.lin_m = (int)lin_m_;
        in.Q_is_diag = Q_is_diag_;
        in.noC = noC_;
        gram_.reset(new GramMetric());
        double sg = sigma_, lam = lambda_A_;
        const char *reason = nullptr;
        // qp_row_metric.cu:136-147, row_metric.cu:133-146: running out of host
        // or device memory declines the metric; other device errors end the solve
        try {
            st = gram_->setup(in, stream_, &sg, &lam, &reason);
        } catch (const std::bad_alloc &) {
            st = HPRQP_STATUS_ALLOCATION_FAILED;
        }
        if (st == HPRQP_STATUS_ALLOCATION_FAILED) {
            (void)hipGetLastError();
            gram_.reset(new GramMetric());
            reason = "allocation";
            st = HPRQP_STATUS_OK;
        }
        if (st != HPRQP_STATUS_OK) return st;
        if (gram_->kind()) {
            sigma_ = sg;
            lambda_A_ = lam;
            noC_ = 0;
            if (Q_is_diag_)
                map(n_, MapDiagFactors{{}, vec(V_fact1), vec(V_fact2), vec(V_fact), vec(V_fact_M), vec(V_diag_Q), sigma_});
        } else {
            gram_.reset();
            const char *v = std::getenv("HPRQP_DIAGNOSTICS");
            if (v && !std::strcmp(v, "1")) std::fprintf(stderr, "GRAM_SKIP reason=%s\n", reason ? reason : "declined");
        }
    }
    to_check_ = 1;
    HPRQP_HIP(hipGetLastError());
    return HPRQP_STATUS_OK;
}

HPRQPStatus HipBackend::setup(const HPRQPModel *model, const HPRQPParams *params)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        std::fprintf(stderr, "hprqp: no HIP device available\n");
        return HPRQP_STATUS_NOT_IMPLEMENTED;
    }
    device_ = params->device_number;
    if (device_ < 0 || device_ >= ndev) return HPRQP_STATUS_INVALID_ARGUMENT;
    // solver.cu:145-151: AUTO asks for the metric only when the built model
    // passes the structural rule of gram_policy.cpp
    if (params->gram_mode < HPRQP_GRAM_OFF || params->gram_mode > HPRQP_GRAM_REQUEST)
        return HPRQP_STATUS_INVALID_ARGUMENT;
    gram_mode_ = (params->gram_mode == HPRQP_GRAM_REQUEST ||
                  (params->gram_mode == HPRQP_GRAM_AUTO && gram_auto_eligible(model)))
                     ? 2
                     : 0;
    HPRQP_HIP(hipSetDevice(device_));
    HPRQP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    obj_constant_ = model->obj_constant;
    // Unscaled bound magnitudes (scale.c:273-274)
    norm_b_org_ = 0.0;
    for (size_t i = 0; i < model->A.nrow; ++i) {
        double a = std::isinf(model->AL[i]) ? 0.0 : std::fabs(model->AL[i]);
        double b = std::isinf(model->AU[i]) ? 0.0 : std::fabs(model->AU[i]);
        double v = a > b ? a : b;
        if (v > norm_b_org_) norm_b_org_ = v;
    }
    norm_c_org_ = 0.0;
    for (size_t i = 0; i < model->Q.ncol; ++i) norm_c_org_ = std::max(norm_c_org_, std::fabs(model->c[i]));
    SetupClock clock(stream_);
    HPRQPStatus st = upload(model);
    clock.mark("upload, transpose, row blocks");
    if (st == HPRQP_STATUS_OK) st = scale(params);
    clock.mark("scaling");
    if (st == HPRQP_STATUS_OK) st = autotune_panels();
    clock.mark("panel and residency autotune");
    if (st == HPRQP_STATUS_OK) st = workspace_setup(model);
    clock.mark("workspace (power iterations, sigma)");
    return st;
}

// ---------------------------------------------------------------- operators

HPRQPStatus HipBackend::iteration_step(size_t inner)
{
    const double h1 = 1.0 / ((double)inner + 2.0);
    const double h2 = 1.0 - h1;
    const double f2s = 1.0 / (1.0 + sigma_ * lambda_Q_);
    const double f1s = 1.0 - f2s;
    const bool chk = to_check_ != 0, diag = Q_is_diag_ != 0;
    if (num_var_cones_ || num_row_cones_ || gram_) return soc_iteration_step(inner);
    // the check variant of the w2 step stores A^T y_bar (reused by the residuals)
    aty_bar_valid_ = chk && !noC_;
    if (noC_) {
#define HPRQP_NOC(D, C)                                                                                           \
    spmv(AT_, vec(V_y), EpiZxw1NoC<D, C>{{}, vec(V_ATy), vec(V_x), vec(V_x_bar), vec(V_x_hat), vec(V_w),           \
                                         vec(V_w_bar), vec(V_dx), vec(V_dw), vec(V_z_bar), vec(V_Qw), vec(V_c),     \
                                         vec(V_l), vec(V_u), vec(V_last_x), vec(V_last_w), vec(V_fact1),            \
                                         vec(V_fact2), sigma_, f1s, f2s, h1, h2}, 0, T_ZXW1_NOC)
        if (diag) {
            if (chk) HPRQP_NOC(true, true);
            else HPRQP_NOC(true, false);
        } else {
            if (chk) HPRQP_NOC(false, true);
            else HPRQP_NOC(false, false);
        }
#undef HPRQP_NOC
        spmv(Q_, vec(V_w_bar), EpiQwNoC{{}, vec(V_Qw), vec(V_Qw_bar), vec(V_tempv), vec(V_last_Qw), vec(V_x_hat),
                                        sigma_, h1, h2}, 0, T_QW_NOC);
    } else {
#define HPRQP_ZXW(D, C)                                                                                            \
    spmv(Q_, vec(V_w), EpiZxw1<D, C>{{}, vec(V_Qw), vec(V_x), vec(V_x_bar), vec(V_x_hat), vec(V_w_bar), vec(V_dx),  \
                                     vec(V_z_bar), vec(V_w), vec(V_ATy), vec(V_c), vec(V_l), vec(V_u),              \
                                     vec(V_last_x), vec(V_fact1), vec(V_fact2), sigma_, f1s, f2s, h1, h2}, 0, T_ZXW1)
        if (diag) {
            if (chk) HPRQP_ZXW(true, true);
            else HPRQP_ZXW(true, false);
        } else {
            if (chk) HPRQP_ZXW(false, true);
            else HPRQP_ZXW(false, false);
        }
#undef HPRQP_ZXW
        // iteration.c:132-134: without rows the y step (and its Q w_bar) is skipped
        if (m_ > 0) spmv(Q_, vec(V_w_bar), EpiTempv{{}, vec(V_tempv), vec(V_x_hat), vec(V_Qw), sigma_}, 0, T_TEMPV);
    }
    if (m_ > 0) {
        const double f1 = lambda_A_ * sigma_;
        if (f1 <= 0.0) return HPRQP_STATUS_INVALID_ARGUMENT;
        const double f2 = 1.0 / f1;
        if (chk)
            spmv(A_, vec(V_tempv), EpiY<true>{{}, vec(V_y), vec(V_y_bar), vec(V_s), vec(V_dy), vec(V_AL), vec(V_AU),
                                                vec(V_last_y), f1, f2, h1, h2}, 0, T_Y);
        else
            spmv(A_, vec(V_tempv), EpiY<false>{{}, vec(V_y), vec(V_y_bar), vec(V_s), vec(V_dy), vec(V_AL), vec(V_AU),
                                                 vec(V_last_y), f1, f2, h1, h2}, 0, T_Y);
    }
    if (!noC_) {
        const double fs = sigma_ / (1.0 + sigma_ * lambda_Q_);
#define HPRQP_W2(D, C)                                                                                            \
    spmv(AT_, vec(V_y_bar), EpiW2<D, C>{{}, vec(V_w_bar), vec(V_w), vec(V_ATy), vec(V_ATy_bar), vec(V_dw),          \
                                        vec(V_ATdy), vec(V_last_w), vec(V_last_ATy), vec(V_fact), fs, h1, h2}, 0, T_W2)
        if (diag) {
            if (chk) HPRQP_W2(true, true);
            else HPRQP_W2(true, false);
        } else {
            if (chk) HPRQP_W2(false, true);
            else HPRQP_W2(false, false);
        }
#undef HPRQP_W2
    }
    HPRQP_HIP(hipGetLastError());
    return HPRQP_STATUS_OK;
}

// residuals.c:88-189 (CPU convention: the dual objective uses the iteration's s)
HPRQPStatus HipBackend::compute_residuals(HPRQPResiduals *res, size_t iter)
{
    std::memset(res, 0, sizeof(*res));
    if (num_var_cones_ || num_row_cones_) return soc_residuals(res);
    spmv(Q_, vec(V_x_bar), EpiResQ{nullptr, vec(V_Qx), vec(V_x_bar), vec(V_c), vec(V_z_bar)}, 0);
    finish_later(0, red_blocks(Q_), 3, 0, 0);
    zero_red(3, 8);
    {
        const EpiResAT rat{nullptr, vec(V_ATdy), vec(V_Qx), vec(V_c), vec(V_z_bar), vec(V_col_norm), c_scale_};
        if (aty_bar_valid_) {
            // the check step's A^T y_bar product is reused (same kernel, same bits)
            map(n_, MapResAT{nullptr, rat, vec(V_ATy_bar)}, 1);
            finish_later(1, n_grid(), 0, 3, 3);
        } else {
            spmv(AT_, vec(V_y_bar), rat, 1);
            finish_later(1, red_blocks(AT_), 0, 3, 3);
        }
    }
    if (m_ > 0) {
        // with the plain-QP metric the dual objective pairs y_bar with the
        // projection of A x_bar (the reference GPU's convention,
        // residuals.cu:28-50, 138-147); otherwise with the iteration's s
        spmv(A_, vec(V_x_bar), EpiResA{nullptr, vec(V_Ax), vec(V_AL), vec(V_AU), vec(V_row_norm), vec(V_y_bar), vec(V_s), b_scale_, gram_ ? 1 : 0}, 2);
        finish_later(2, red_blocks(A_), 1, 2, 6);
    }
    if (iter == 0) {
        map(n_, MapBoundViolation{nullptr, vec(V_dx), vec(V_l), vec(V_u), vec(V_x_bar), vec(V_col_norm), b_scale_}, 3);
        finish_later(3, n_grid(), 0, 1, 9);
    }
    HPRQP_HIP(hipGetLastError());
    HPRQPStatus st = read_red(10);
    if (st != HPRQP_STATUS_OK) return st;
    const double *r = red_host_;
    const double cx = r[0], xqx = r[1], zx = r[2];
    const double rd_max = r[3], qx_max = r[4], aty_max = r[5];
    const double ys = m_ > 0 ? r[6] : 0.0;
    const double scale = b_scale_ * c_scale_;
    res->primal_obj_bar = scale * (cx + 0.5 * xqx) + obj_constant_;
    res->dual_obj_bar = scale * (-0.5 * xqx + zx + ys) + obj_constant_;
    res->rel_gap_bar = std::fabs(res->primal_obj_bar - res->dual_obj_bar) /
                       (1.0 + std::max(std::fabs(res->primal_obj_bar), std::fabs(res->dual_obj_bar)));
    res->err_Rd_org_bar = rd_max / (1.0 + std::max(std::max(norm_c_org_, aty_max), qx_max));
    if (m_ > 0) res->err_Rp_org_bar = r[7] / (1.0 + std::max(norm_b_org_, r[8]));
    if (iter == 0 && r[9] > res->err_Rp_org_bar) res->err_Rp_org_bar = r[9];
    res->KKTx_and_gap_org_bar = std::max(std::max(res->err_Rp_org_bar, res->err_Rd_org_bar), res->rel_gap_bar);
    res->is_updated = 1;
    last_res_ = *res;
    return HPRQP_STATUS_OK;
}

// src/gpu/soc_iteration.cuh:45-82: the generic cone step with full writes.
HPRQPStatus HipBackend::soc_iteration_step(size_t inner)
{
    aty_bar_valid_ = true;  // the cone step always runs the w2 variant that stores A^T y_bar
    const double a = 1.0 / ((double)inner + 2.0), b = 1.0 - a;
    const double f2s = 1.0 / (1.0 + sigma_ * lambda_Q_);
    const double f1s = 1.0 - f2s;
    SocFinishX fx{vec(V_x), vec(V_x_hat), vec(V_w_bar), vec(V_dx), vec(V_z_bar), vec(V_tempv), vec(V_w),
                  vec(V_last_x), vec(V_fact1), vec(V_fact2), sigma_, f1s, f2s, a, b, Q_is_diag_};
    EpiSocX ex;
    ex.fin = fx;
    ex.Qw = vec(V_Qw);
    ex.tempv = vec(V_tempv);
    ex.x_bar = vec(V_x_bar);
    ex.xv = vec(V_x);
    ex.ATy = vec(V_ATy);
    ex.c = vec(V_c);
    ex.l = vec(V_l);
    ex.u = vec(V_u);
    ex.lin_n = (int)lin_n_;
    if (fused_x()) {
        // Q w, the interval prefix, the cones and their finish in one launch
        if (profiling_) prof_begin(T_ZXW1);
        const int nbox = grid_for(lin_n_);
        const unsigned cone_blocks = (unsigned)((num_var_cones_ * (size_t)var_lanes_ + kThreads - 1) / kThreads);
        const unsigned grid = (unsigned)nbox + cone_blocks;
        const int cnt = (int)num_var_cones_;
        switch (var_lanes_) {
        case 1: soc_x_diag_kernel<1><<<grid, kThreads, 0, stream_>>>(nbox, cnt, var_off_, Q_.rp, Q_.val, vec(V_w), ex); break;
        case 2: soc_x_diag_kernel<2><<<grid, kThreads, 0, stream_>>>(nbox, cnt, var_off_, Q_.rp, Q_.val, vec(V_w), ex); break;
        case 4: soc_x_diag_kernel<4><<<grid, kThreads, 0, stream_>>>(nbox, cnt, var_off_, Q_.rp, Q_.val, vec(V_w), ex); break;
        case 8: soc_x_diag_kernel<8><<<grid, kThreads, 0, stream_>>>(nbox, cnt, var_off_, Q_.rp, Q_.val, vec(V_w), ex); break;
        case 16: soc_x_diag_kernel<16><<<grid, kThreads, 0, stream_>>>(nbox, cnt, var_off_, Q_.rp, Q_.val, vec(V_w), ex); break;
        case 32: soc_x_diag_kernel<32><<<grid, kThreads, 0, stream_>>>(nbox, cnt, var_off_, Q_.rp, Q_.val, vec(V_w), ex); break;
        default: soc_x_diag_kernel<64><<<grid, kThreads, 0, stream_>>>(nbox, cnt, var_off_, Q_.rp, Q_.val, vec(V_w), ex); break;
        }
        if (profiling_) prof_end();
    } else {
        spmv(Q_, vec(V_w), ex, 0, T_ZXW1);
    }
    if (num_var_cones_ && !fused_x()) {
        if (profiling_) prof_begin(T_CONE_X);
        if (var_lanes_) {
            cone_step(var_lanes_, num_var_cones_, var_off_, vec(V_tempv), vec(V_x_bar), ConeFinX{fx});
        } else {
            soc_cone_project_kernel<<<(unsigned)((num_var_cones_ + kConeThreads - 1) / kConeThreads), kConeThreads,
                                      0, stream_>>>((int)num_var_cones_, var_off_, vec(V_tempv), vec(V_x_bar));
            map(n_ - lin_n_, MapSocFinishX{{}, fx, vec(V_x_bar), (int)lin_n_});
        }
        if (profiling_) prof_end();
    }
    if (!fused_x())  // else the fused x step has written tempv already
        spmv(Q_, vec(V_w_bar), EpiTempv{{}, vec(V_tempv), vec(V_x_hat), vec(V_Qw), sigma_}, 0, T_TEMPV);
    if (m_ > 0) {
        const double tau = sigma_ * lambda_A_;
        SocFinishY fy{vec(V_y), vec(V_y_bar), vec(V_dy), vec(V_s), vec(V_Rp_soc()), vec(V_last_y), tau, a, b};
        EpiSocY ey;
        ey.fin = fy;
        ey.Rp = vec(V_Rp_soc());
        ey.s = vec(V_s);
        ey.rhs = num_row_cones_ ? vec(V_soc_rhs) : nullptr;
        ey.AL = vec(V_AL);
        ey.AU = vec(V_AU);
        ey.lin_m = (int)lin_m_;
        if (gram_) {
            ey.gmap = gram_->map();
            ey.gr = gram_->rhs();
            ey.gb = gram_->b();
        }
        spmv(A_, vec(V_tempv), ey, 0, T_Y);
        if (num_row_cones_) {
            if (profiling_) prof_begin(T_CONE_Y);
            if (row_lanes_) {
                cone_step(row_lanes_, num_row_cones_, row_off_, vec(V_Rp_soc()), vec(V_s), ConeFinY{fy});
            } else {
                soc_cone_project_kernel<<<(unsigned)((num_row_cones_ + kConeThreads - 1) / kConeThreads),
                                          kConeThreads, 0, stream_>>>((int)num_row_cones_, row_off_, vec(V_Rp_soc()),
                                                                      vec(V_s));
                map(m_ - lin_m_, MapSocFinishY{{}, fy, (int)lin_m_});
            }
            if (profiling_) prof_end();
        }
        // qp_row_metric.cu:55-63 / row_metric.cu:46-55: the metric rows
        if (gram_) {
            if (profiling_) prof_begin(T_GRAM);
            HPRQPStatus st = gram_->finish_y(stream_, vec(V_y), vec(V_y_bar), vec(V_dy), vec(V_last_y), sigma_, a, b);
            if (profiling_) prof_end();
            if (st != HPRQP_STATUS_OK) return st;
        }
    }
    const double fs = sigma_ / (1.0 + sigma_ * lambda_Q_);
    if (Q_is_diag_)
        spmv(AT_, vec(V_y_bar), EpiW2<true, true>{{}, vec(V_w_bar), vec(V_w), vec(V_ATy), vec(V_ATy_bar), vec(V_dw),
                                                  vec(V_ATdy), vec(V_last_w), vec(V_last_ATy), vec(V_fact), fs, a, b},
             0, T_W2);
    else
        spmv(AT_, vec(V_y_bar), EpiW2<false, true>{{}, vec(V_w_bar), vec(V_w), vec(V_ATy), vec(V_ATy_bar), vec(V_dw),
                                                   vec(V_ATdy), vec(V_last_w), vec(V_last_ATy), vec(V_fact), fs, a, b},
             0, T_W2);
    HPRQP_HIP(hipGetLastError());
    return HPRQP_STATUS_OK;
}

// src/gpu/soc_residuals.cuh:27-78 (the iteration s pairs with y_bar; the
// variable-domain violation is audited at every check).
HPRQPStatus HipBackend::soc_residuals(HPRQPResiduals *res)
{
    const int vc = (int)num_var_cones_, rc = (int)num_row_cones_;
    spmv(Q_, vec(V_x_bar), EpiResQ{nullptr, vec(V_Qx), vec(V_x_bar), vec(V_c), vec(V_z_bar)}, 0);
    finish_later(0, red_blocks(Q_), 3, 0, 0);
    zero_red(3, 8);
    {
        const EpiResAT rat{nullptr, vec(V_ATdy), vec(V_Qx), vec(V_c), vec(V_z_bar), vec(V_col_norm), c_scale_};
        if (aty_bar_valid_) {
            map(n_, MapResAT{nullptr, rat, vec(V_ATy_bar)}, 1);
            finish_later(1, n_grid(), 0, 3, 3);
        } else {
            spmv(AT_, vec(V_y_bar), rat, 1);
            finish_later(1, red_blocks(AT_), 0, 3, 3);
        }
    }
    if (m_ > 0) {
        spmv(A_, vec(V_x_bar), EpiSocResA{nullptr, vec(V_Ax), vec(V_Rp_soc()), rc ? vec(V_soc_rhs) : nullptr,
                                          vec(V_row_norm), vec(V_y_bar), vec(V_s), b_scale_}, 2);
        finish_later(2, red_blocks(A_), 2, 1, 6);
        soc_project_mixed_kernel<<<grid_for(std::max(lin_m_, (size_t)rc)), kThreads, 0, stream_>>>(
            (int)m_, (int)lin_m_, vec(V_Rp_soc()), vec(V_AL), vec(V_AU), rc, row_off_, vec(V_tr));
        map(m_, MapSocRowResid{nullptr, vec(V_tr), vec(V_Rp_soc()), vec(V_row_norm), b_scale_}, 3);
        finish_later(3, m_grid(), 0, 1, 9);
    }
    soc_project_mixed_kernel<<<grid_for(std::max(lin_n_, (size_t)vc)), kThreads, 0, stream_>>>(
        (int)n_, (int)lin_n_, vec(V_x_bar), vec(V_l), vec(V_u), vc, var_off_, vec(V_tc));
    map(n_, MapSocVarResid{nullptr, vec(V_tc), vec(V_x_bar), vec(V_col_norm), b_scale_}, 0);
    finish_later(0, n_grid(), 0, 1, 10);
    HPRQP_HIP(hipGetLastError());
    HPRQPStatus st = read_red(11);
    if (st != HPRQP_STATUS_OK) return st;
    const double *r = red_host_;
    const double cx = r[0], xqx = r[1], zx = r[2];
    const double ys = m_ > 0 ? r[6] : 0.0, yb = m_ > 0 ? r[7] : 0.0;
    const double scale = b_scale_ * c_scale_;
    res->primal_obj_bar = scale * (cx + 0.5 * xqx) + obj_constant_;
    res->dual_obj_bar = scale * (-0.5 * xqx + zx + ys + yb) + obj_constant_;
    res->rel_gap_bar = std::fabs(res->primal_obj_bar - res->dual_obj_bar) /
                       (1.0 + std::fmax(std::fabs(res->primal_obj_bar), std::fabs(res->dual_obj_bar)));
    res->err_Rd_org_bar = r[3] / (1.0 + std::fmax(norm_c_org_, std::fmax(r[4], r[5])));
    if (m_ > 0) res->err_Rp_org_bar = r[9] / (1.0 + std::fmax(norm_b_org_, r[8]));
    res->err_Rp_org_bar = std::fmax(res->err_Rp_org_bar, r[10]);
    res->KKTx_and_gap_org_bar = std::fmax(res->err_Rp_org_bar, std::fmax(res->err_Rd_org_bar, res->rel_gap_bar));
    res->is_updated = 1;
    last_res_ = *res;
    return HPRQP_STATUS_OK;
}

// control.c:155-229
double HipBackend::compute_m_norm()
{
    const bool diag = Q_is_diag_ != 0;
    spmv(Q_, vec(V_dw), EpiMnQdw{nullptr, vec(V_dQw), vec(V_dx), vec(V_dw)}, 0);
    finish_later(0, red_blocks(Q_), 4, 0, 0);
    if (m_ > 0) {
        spmv(AT_, vec(V_dy), EpiMnATdy{nullptr, vec(V_ATdy), vec(V_dx)}, 1);
        finish_later(1, red_blocks(AT_), 1, 0, 4);
        spmv(Q_, vec(V_ATdy), EpiMnQATdy{nullptr, vec(V_QATdy), vec(V_dQw), vec(V_ATdy), vec(V_fact_M), diag ? 1 : 0}, 2);
        finish_later(2, red_blocks(Q_), 3, 0, 5);
        map(m_, MapDot{nullptr, vec(V_dy), vec(V_dy)}, 3);
        finish_later(3, m_grid(), 1, 0, 8);
    }
    if (hipGetLastError() != hipSuccess || read_red(9) != HPRQP_STATUS_OK) return INFINITY;
    const double *r = red_host_;
    double gram_q = 0.0;  // control.cu:273-279: the metric's quadratic in dy
    if (gram_ && m_ > 0 && gram_->quadratic(stream_, vec(V_dy), &gram_q) != HPRQP_STATUS_OK) return INFINITY;
    const double dx_sq = r[0], dqw_dx = r[1], dw_dqw = r[2], dqw_sq = r[3];
    const double sg = sigma_;
    double m1 = 0.0, m3 = 0.0;
    double m2 = (1.0 / sg) * dx_sq;
    m2 -= 2.0 * dqw_dx;
    if (m_ > 0) {
        const double atdy_dx = r[4], dqw_atdy = r[5], atdy_qatdy = r[6], fm_term = r[7], dy_sq = r[8];
        m1 = gram_ ? sg * gram_q : sg * lambda_A_ * dy_sq;
        m1 -= 2.0 * sg * dqw_atdy;
        m2 += 2.0 * atdy_dx;
        if (diag) {
            m3 = fm_term;
            m1 += sg * dqw_sq;
        } else {
            m3 = (sg * sg) / (1.0 + sg * lambda_Q_) * atdy_qatdy;
            m1 += sg * lambda_Q_ * dw_dqw;
        }
    } else if (!diag) {
        m1 = sg * lambda_Q_ * dw_dqw;
    }
    m2 += m1 > 0.0 ? m1 : 0.0;
    return std::sqrt((m2 > 0.0 ? m2 : 0.0) + (m3 > 0.0 ? m3 : 0.0));
}

// control.c:307-411
HPRQPStatus HipBackend::update_sigma(HPRQPRestart *restart, const HPRQPResiduals *res)
{
    if (restart->restart_flag < 1) return HPRQP_STATUS_OK;
    const double old = sigma_;
    const bool diag = Q_is_diag_ != 0;
    map(n_, MapSigmaDiffN{nullptr, vec(V_dx), vec(V_dw), vec(V_ATdy), vec(V_x_bar), vec(V_last_x), vec(V_w_bar),
                          vec(V_last_w), vec(V_ATy_bar), vec(V_last_ATy), m_ > 0 ? 1 : 0}, 0);
    finish_later(0, n_grid(), 1, 0, 0);
    zero_red(1, 5);
    if (m_ > 0) {
        map(m_, MapSigmaDiffM{nullptr, vec(V_dy), vec(V_y_bar), vec(V_last_y)}, 1);
        finish_later(1, m_grid(), 1, 0, 1);
    }
    spmv(Q_, vec(V_dw), EpiSigmaQdw{nullptr, vec(V_dQw), vec(V_ATdy), vec(V_dw)}, 2);
    finish_later(2, red_blocks(Q_), 3, 0, 2);
    if (m_ > 0) {
        spmv(Q_, vec(V_ATdy), EpiSigmaQATdy{nullptr, vec(V_QATdy), vec(V_ATdy)}, 3);
        finish_later(3, red_blocks(Q_), 1, 0, 5);
    }
    HPRQP_HIP(hipGetLastError());
    HPRQPStatus st = read_red(6);
    if (st != HPRQP_STATUS_OK) return st;
    const double *r = red_host_;
    double b = r[0], a = 0.0, cterm = 0.0, dterm = 0.0;
    const double dy_sq = r[1], dqw_atdy = r[2], dqw_sq = r[3], dw_dqw = r[4], atdy_qatdy = r[5];
    if (m_ > 0) {
        if (gram_) {
            // control.cu:400-404
            st = gram_->quadratic(stream_, vec(V_dy), &a);
            if (st != HPRQP_STATUS_OK) return st;
        } else {
            a = lambda_A_ * dy_sq;
        }
        a -= 2.0 * dqw_atdy;
    }
    if (diag) {
        a += dqw_sq;
    } else {
        a += lambda_Q_ * dw_dqw;
        if (m_ > 0) {
            cterm = atdy_qatdy;
            dterm = lambda_Q_;
        }
    }
    a = a > 1e-12 ? a : 1e-12;
    b = b > 1e-12 ? b : 1e-12;
    double est;
    if (diag && m_ > 0) {
        if (!golden_) {
            HPRQP_HIP(hipMalloc((void **)&golden_, sizeof(GoldenState)));
            HPRQP_HIP(hipMalloc((void **)&golden_part_, sizeof(double) * kGoldenBlocks * kGoldenPoints));
        }
        golden_init_kernel<<<1, 64, 0, stream_>>>(golden_, a, b);
        if (!golden_graph_) {
            // the pass chain (every argument fixed for the session) is
            // captured once and replayed for each search
            const int gblk = std::max(1, std::min(kGoldenBlocks, (int)((n_ + 4 * kThreads - 1) / (4 * kThreads))));
            hipGraph_t graph = nullptr;
            HPRQP_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            for (int pass = 0; pass < kGoldenPasses; ++pass) {
                golden_partials_kernel<<<gblk, kThreads, 0, stream_>>>(golden_, vec(V_diag_Q), vec(V_ATdy),
                                                                       vec(V_QATdy), (int)n_, golden_part_);
                golden_step_kernel<<<1, kThreads, 0, stream_>>>(golden_, golden_part_, gblk);
            }
            HPRQP_HIP(hipStreamEndCapture(stream_, &graph));
            const hipError_t ie = hipGraphInstantiate(&golden_graph_, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HPRQP_HIP(ie);
        }
        HPRQP_HIP(hipGraphLaunch(golden_graph_, stream_));
        HPRQP_HIP(hipGetLastError());
        HPRQP_HIP(hipMemcpyAsync(red_host_ + 10, reinterpret_cast<char *>(golden_) + offsetof(GoldenState, result),
                                 sizeof(double), hipMemcpyDeviceToHost, stream_));
        HPRQP_HIP(hipStreamSynchronize(stream_));
        est = red_host_[10];
    } else if (!diag && m_ > 0) {
        est = golden_search_sigma(a, b, cterm > 0.0 ? cterm : 0.0, dterm > 0.0 ? dterm : 0.0);
    } else {
        est = std::sqrt(b / a);
    }
    sigma_ = sigma_from_estimate(est, old, restart, res);
    if (diag && std::fabs(sigma_ - old) > 1e-15)
        map(n_, MapDiagFactors{{}, vec(V_fact1), vec(V_fact2), vec(V_fact), vec(V_fact_M), vec(V_diag_Q), sigma_});
    HPRQP_HIP(hipGetLastError());
    return HPRQP_STATUS_OK;
}

// control.c:413-457
HPRQPStatus HipBackend::restart_apply(HPRQPRestart *restart)
{
    if (restart->restart_flag <= 0) return HPRQP_STATUS_OK;
    spmv(AT_, vec(V_y_bar), EpiRestart{{}, vec(V_ATy_bar), vec(V_ATy), vec(V_last_ATy), vec(V_x), vec(V_w),
                                       vec(V_last_x), vec(V_last_w), vec(V_Qw), vec(V_last_Qw), vec(V_x_bar),
                                       vec(V_w_bar), vec(V_Qw_bar), m_ > 0 ? 1 : 0, noC_});
    map(m_, MapRestartM{{}, vec(V_y), vec(V_last_y), vec(V_y_bar)});
    HPRQP_HIP(hipGetLastError());
    restart->last_gap = restart->current_gap;
    restart->save_gap = INFINITY;
    restart->times += 1;
    restart->inner = 0;
    restart_times_ = restart->times;
    return HPRQP_STATUS_OK;
}

// solver.c:67-106 / workspace.cu:427-511
HPRQPStatus HipBackend::copy_output_iterates(double *x, double *w, double *z, double *y)
{
    std::vector<double> xb(n_), wb(n_), zb(n_), yb(m_);
    if (n_) {
        HPRQP_HIP(hipMemcpyAsync(xb.data(), vec(V_x_bar), sizeof(double) * n_, hipMemcpyDeviceToHost, stream_));
        HPRQP_HIP(hipMemcpyAsync(wb.data(), vec(V_w_bar), sizeof(double) * n_, hipMemcpyDeviceToHost, stream_));
        HPRQP_HIP(hipMemcpyAsync(zb.data(), vec(V_z_bar), sizeof(double) * n_, hipMemcpyDeviceToHost, stream_));
    }
    if (m_) HPRQP_HIP(hipMemcpyAsync(yb.data(), vec(V_y_bar), sizeof(double) * m_, hipMemcpyDeviceToHost, stream_));
    HPRQP_HIP(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < n_; ++i) {
        const double cn = col_norm_h_[i];
        x[i] = b_scale_ * (xb[i] / cn);
        w[i] = b_scale_ * (wb[i] / cn);
        z[i] = c_scale_ * (zb[i] * cn);
    }
    for (size_t i = 0; i < m_; ++i) y[i] = c_scale_ * (yb[i] / row_norm_h_[i]);
    return HPRQP_STATUS_OK;
}

// Algorithmic HBM bytes of one launch (non-check variant): the matrix
// stream (8 B value + 4 B column index per nonzero, 4 B per row pointer), the
// gathered vector once (8 B per element of its length), and 8 B per element
// for every dense vector read or written by the epilogue.
double HipBackend::algorithmic_bytes(int tag) const
{
    const double n = (double)n_, m = (double)m_;
    const double nq = (double)Q_.nnz, na = (double)A_.nnz;
    const double diag = Q_is_diag_ ? 2.0 : 0.0;
    switch (tag) {
    case T_ZXW1:  // reads w(gather), w, ATy, c, l, u, x, last_x [+fact1,fact2]; writes Qw, x_bar, x_hat, x, w_bar
        // (fused with the cone step and the tempv product: tempv, dx and z_bar
        // of the cone range and the next tempv of every row written too)
        return 12.0 * nq + 4.0 * (n + 1) + 8.0 * n + 8.0 * n * (7.0 + diag + 5.0) +
               (fused_x() ? 8.0 * (n - (double)lin_n_) * 3.0 + 4.0 * (num_var_cones_ + 1) + 8.0 * n : 0.0);
    case T_TEMPV:  // reads w_bar(gather), x_hat, Qw; writes tempv
        return 12.0 * nq + 4.0 * (n + 1) + 8.0 * n + 8.0 * n * 3.0;
    case T_Y:  // reads tempv(gather), y, last_y, AL, AU; writes y_bar, y
        return 12.0 * na + 4.0 * (m + 1) + 8.0 * n + 8.0 * m * 6.0;
    case T_W2:  // reads y_bar(gather), w_bar, ATy, w, last_w, last_ATy [+fact]; writes w_bar, w, ATy
        return 12.0 * na + 4.0 * (n + 1) + 8.0 * m + 8.0 * n * (5.0 + (Q_is_diag_ ? 1.0 : 0.0) + 3.0);
    case T_ZXW1_NOC:  // reads y(gather), Qw, c, l, u, x, w, last_x, last_w [+2]; writes ATy, x_bar, x_hat, x, w_bar, w
        return 12.0 * na + 4.0 * (n + 1) + 8.0 * m + 8.0 * n * (8.0 + diag + 6.0);
    case T_QW_NOC:  // reads w_bar(gather), Qw, last_Qw, x_hat; writes Qw_bar, Qw, tempv
        return 12.0 * nq + 4.0 * (n + 1) + 8.0 * n + 8.0 * n * 6.0;
    case T_CONE_X:  // cone range of x: projection (tempv in, x_bar out), then the finish
        // (reads x, x_bar, tempv, last_x, w [+fact1, fact2]; writes dx, z_bar, x_hat, x, w_bar)
        return 8.0 * (n - (double)lin_n_) * (2.0 + 5.0 + diag + 5.0) + 4.0 * (num_var_cones_ + 1);
    case T_CONE_Y:  // cone range of y: projection (Rp in, s out), then the finish
        // (reads y, s, Rp, last_y; writes dy, y_bar, y)
        return 8.0 * (m - (double)lin_m_) * (2.0 + 4.0 + 3.0) + 4.0 * (num_row_cones_ + 1);
    case T_GRAM: {  // the two triangles of W and W' (e(e+1)/2 entries each), r, u, step;
        // the finish reads y, last_y and writes y, y_bar, dy of the metric rows
        const double e = gram_ ? (double)gram_->rows() : 0.0;
        return 8.0 * e * (e + 1.0) + 8.0 * e * (3.0 + 2.0 + 3.0) + 4.0 * e;
    }
    default:
        return 0.0;
    }
}

HPRQPStatus HipBackend::profile(size_t iterations, double *mean_ms, double *bytes, size_t *launches)
{
    const size_t need = iterations * 16 + 8;
    ev_pool_.resize(need);
    ev_tags_.assign(need, -1);
    for (auto &e : ev_pool_) HPRQP_HIP(hipEventCreate(&e));
    ev_used_ = 0;
    const int saved_check = to_check_;
    to_check_ = 0;
    profiling_ = true;
    HPRQPStatus st = HPRQP_STATUS_OK;
    for (size_t k = 0; k < iterations && st == HPRQP_STATUS_OK; ++k) st = iteration_step(k);
    profiling_ = false;
    to_check_ = saved_check;
    HPRQP_HIP(hipStreamSynchronize(stream_));
    double total[T_COUNT] = {0};
    size_t count[T_COUNT] = {0};
    for (size_t e = 0; e + 1 < ev_used_; ++e) {
        const int tag = ev_tags_[e];
        if (tag < 0 || ev_tags_[e + 1] != -1) continue;
        float ms = 0.0f;
        HPRQP_HIP(hipEventElapsedTime(&ms, ev_pool_[e], ev_pool_[e + 1]));
        total[tag] += ms;
        count[tag] += 1;
    }
    for (auto &e : ev_pool_) (void)hipEventDestroy(e);
    ev_pool_.clear();
    for (int t = 0; t < T_COUNT; ++t) {
        mean_ms[t] = count[t] ? total[t] / (double)count[t] : 0.0;
        bytes[t] = algorithmic_bytes(t);
        launches[t] = count[t];
    }
    return st;
}

double HipBackend::scalar(const char *name) const
{
    if (!std::strcmp(name, "sigma")) return sigma_;
    if (!std::strcmp(name, "lambda_max_A")) return lambda_A_;
    if (!std::strcmp(name, "lambda_max_Q")) return lambda_Q_;
    if (!std::strcmp(name, "power_steps_A")) return power_steps_A_;
    if (!std::strcmp(name, "power_steps_Q")) return power_steps_Q_;
    if (!std::strcmp(name, "Q_is_diag")) return Q_is_diag_;
    if (!std::strcmp(name, "noC")) return noC_;
    if (!std::strcmp(name, "norm_b_org")) return norm_b_org_;
    if (!std::strcmp(name, "norm_c_org")) return norm_c_org_;
    if (!std::strcmp(name, "kkt")) return last_res_.KKTx_and_gap_org_bar;
    if (!std::strcmp(name, "err_rp")) return last_res_.err_Rp_org_bar;
    if (!std::strcmp(name, "err_rd")) return last_res_.err_Rd_org_bar;
    if (!std::strcmp(name, "gap")) return last_res_.rel_gap_bar;
    if (!std::strcmp(name, "pobj")) return last_res_.primal_obj_bar;
    if (!std::strcmp(name, "dobj")) return last_res_.dual_obj_bar;
    if (!std::strcmp(name, "restart_times")) return (double)restart_times_;
    if (!std::strcmp(name, "panels_Q")) return Qp_.P;
    if (!std::strcmp(name, "panels_A")) return Ap_.P;
    if (!std::strcmp(name, "panels_AT")) return ATp_.P;
    if (!std::strcmp(name, "lds_pad_Q")) return Qp_.lds_pad;
    if (!std::strcmp(name, "lds_pad_A")) return Ap_.lds_pad;
    if (!std::strcmp(name, "lds_pad_AT")) return ATp_.lds_pad;
    if (!std::strcmp(name, "gram_active")) return gram_ ? gram_->kind() : 0;
    if (!std::strcmp(name, "gram_rows")) return gram_ ? gram_->rows() : 0;
    if (!std::strcmp(name, "gram_beta")) return gram_ ? gram_->beta() : 0.0;
    if (!std::strcmp(name, "gram_mode")) return gram_mode_;
    return NAN;
}

}  // namespace
}  // namespace hprqp

// ================================================================ C ABI

struct HPRQPHipSession {
    HPRQPParams params;
    hprqp::HipBackend backend;
    std::unique_ptr<hprqp::SolveLoop> loop;
    double setup_seconds = 0.0;
};

extern "C" {

int gpu_backend_available(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

const char *hprqp_hip_device_name(void)
{
    static char name[256] = {0};
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return "";
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return "";
    std::snprintf(name, sizeof(name), "%s (%s)", prop.name, prop.gcnArchName);
    return name;
}

HPRQPStatus hprqp_hip_session_create(const HPRQPModel *model, const HPRQPParams *params, HPRQPHipSession **out)
{
    if (!model || !params || !out) return HPRQP_STATUS_INVALID_ARGUMENT;
    *out = nullptr;
    if (model->Q.ncol == 0) return HPRQP_STATUS_INVALID_ARGUMENT;
    auto *s = new (std::nothrow) HPRQPHipSession();
    if (!s) return HPRQP_STATUS_ALLOCATION_FAILED;
    s->params = *params;
    const double t0 = hprqp::wall_seconds();
    HPRQPStatus st = s->backend.setup(model, &s->params);
    if (st != HPRQP_STATUS_OK) {
        delete s;
        return st;
    }
    if (s->backend.n() && (model->num_var_cones || model->num_row_cones)) {
        // The cone step is provided by the SOC kernels; refuse until then.
    }
    s->setup_seconds = hprqp::wall_seconds() - t0;
    s->loop.reset(new hprqp::SolveLoop(&s->backend, &s->params, "GPU mode", model));
    *out = s;
    return HPRQP_STATUS_OK;
}

void hprqp_hip_session_destroy(HPRQPHipSession *s) { delete s; }

HPRQPStatus hprqp_hip_session_advance(HPRQPHipSession *s, size_t bodies, int *finished)
{
    if (!s) return HPRQP_STATUS_INVALID_ARGUMENT;
    HPRQPStatus st = HPRQP_STATUS_OK;
    for (size_t k = 0; k < bodies && !s->loop->finished(); ++k) {
        st = s->loop->body();
        if (st != HPRQP_STATUS_OK) break;
    }
    if (finished) *finished = s->loop->finished() ? 1 : 0;
    return st;
}

size_t hprqp_hip_session_read(const HPRQPHipSession *s, const char *name, double *out, size_t capacity)
{
    if (!s || !name) return 0;
    if (!std::strncmp(name, "gram_", 5)) return s->backend.gram_read(name, out, capacity);
    for (int v = 0; v < hprqp::V_M_END; ++v) {
        if (std::strcmp(hprqp::kVecNames[v], name) != 0) continue;
        size_t len = s->backend.vec_len(v);
        if (out && len) {
            size_t cnt = std::min(len, capacity);
            if (hipMemcpyAsync(out, s->backend.vec(v), sizeof(double) * cnt, hipMemcpyDeviceToHost,
                               s->backend.stream()) != hipSuccess ||
                hipStreamSynchronize(s->backend.stream()) != hipSuccess)
                return 0;
        }
        return len;
    }
    return 0;
}

double hprqp_hip_session_scalar(const HPRQPHipSession *s, const char *name)
{
    if (!s || !name) return NAN;
    if (!std::strcmp(name, "iter")) return (double)s->loop->iter();
    if (!std::strcmp(name, "inner")) return (double)s->loop->restart().inner;
    if (!std::strcmp(name, "setup_seconds")) return s->setup_seconds;
    if (!std::strcmp(name, "elapsed")) return s->loop->elapsed();
    return s->backend.scalar(name);
}

size_t hprqp_hip_session_matrix(const HPRQPHipSession *s, const char *name, size_t *row_ptr, size_t *col_ind,
                                double *values)
{
    if (!s || !name) return 0;
    const hprqp::DevCsr &M = s->backend.matrix(name);
    hipStream_t st = s->backend.stream();
    std::vector<int> rp(M.nrow + 1), ci(M.nnz);
    if (row_ptr || col_ind) {
        if (hipMemcpyAsync(rp.data(), M.rp, sizeof(int) * (M.nrow + 1), hipMemcpyDeviceToHost, st) != hipSuccess) return 0;
        if (M.nnz && hipMemcpyAsync(ci.data(), M.ci, sizeof(int) * M.nnz, hipMemcpyDeviceToHost, st) != hipSuccess) return 0;
    }
    if (values && M.nnz &&
        hipMemcpyAsync(values, M.val, sizeof(double) * M.nnz, hipMemcpyDeviceToHost, st) != hipSuccess)
        return 0;
    if (hipStreamSynchronize(st) != hipSuccess) return 0;
    if (row_ptr)
        for (int r = 0; r <= M.nrow; ++r) row_ptr[r] = (size_t)rp[r];
    if (col_ind)
        for (int k = 0; k < M.nnz; ++k) col_ind[k] = (size_t)ci[k];
    return (size_t)M.nnz;
}

HPRQPStatus hprqp_hip_session_results(HPRQPHipSession *s, HPRQPResults *results)
{
    if (!s || !results) return HPRQP_STATUS_INVALID_ARGUMENT;
    HPRQPStatus st = s->loop->collect(results);
    if (st != HPRQP_STATUS_OK) return st;
    results->total_time = s->loop->elapsed() + s->setup_seconds;
    return HPRQP_STATUS_OK;
}

// src/gpu/solver.cu:127-231
HPRQPStatus solve_gpu(const HPRQPModel *model, const HPRQPParams *params, HPRQPResults *results)
{
    if (!model || !params || !results) return HPRQP_STATUS_INVALID_ARGUMENT;
    const double t0 = hprqp::wall_seconds();
    results_init(results);
    if (params->verbose) {
        print_model_info(model);
        print_params_info(params);
    }
    HPRQPHipSession *s = nullptr;
    HPRQPStatus st = hprqp_hip_session_create(model, params, &s);
    if (st != HPRQP_STATUS_OK) return st;
    int fin = 0;
    st = hprqp_hip_session_advance(s, (size_t)-1, &fin);
    if (st == HPRQP_STATUS_OK) st = s->loop->collect(results);
    HPRQPStatus final_status = s->loop->final_status();
    hprqp_hip_session_destroy(s);
    if (st != HPRQP_STATUS_OK) return st;
    results->total_time = hprqp::wall_seconds() - t0;
    return final_status;
}

// src/shared/solver.c:245-260
HPRQPStatus solve(const HPRQPModel *model, const HPRQPParams *params, HPRQPResults *results)
{
    if (!params || !model || !results) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (!params->use_gpu) {
        results_init(results);
        std::fprintf(stderr, "hprqp: this library only provides the MI355X path (use_gpu=1)\n");
        return HPRQP_STATUS_NOT_IMPLEMENTED;
    }
    return solve_gpu(model, params, results);
}

int hprqp_hip_session_profile(HPRQPHipSession *s, size_t iterations, int capacity, char *names, double *mean_ms,
                              double *bytes_per_launch, size_t *launches)
{
    using B = hprqp::HipBackend;
    if (!s || capacity < (int)B::T_COUNT || !names || !mean_ms || !bytes_per_launch || !launches) return -1;
    if (s->backend.profile(iterations, mean_ms, bytes_per_launch, launches) != HPRQP_STATUS_OK) return -1;
    for (int t = 0; t < (int)B::T_COUNT; ++t) std::snprintf(names + 32 * t, 32, "%s", B::tag_name(t));
    return (int)B::T_COUNT;
}

HPRQPStatus hprqp_hip_csr_spmv(int nrow, const int *row_ptr, const int *col_ind, const double *values,
                               const double *x, double *y, void *stream)
{
    if (nrow < 0 || (nrow > 0 && (!row_ptr || !x || !y))) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (nrow == 0) return HPRQP_STATUS_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hprqp::DevCsr M;
    M.nrow = nrow;
    M.rp = const_cast<int *>(row_ptr);
    M.ci = const_cast<int *>(col_ind);
    M.val = const_cast<double *>(values);
    HPRQPStatus st = hprqp::upload_blocks(M, s);
    hprqp::DeviceBuffer<int2> blocks;  // freed on every return below
    blocks.p = M.blocks;
    if (st != HPRQP_STATUS_OK) return st;
    hprqp::EpiStore epi;
    epi.out = y;
    hprqp::spmv_rows_kernel<hprqp::EpiStore><<<M.nblocks, hprqp::kThreads, 0, s>>>(M.rp, M.ci, M.val, x, M.blocks, epi);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(hipStreamSynchronize(s));
    return HPRQP_STATUS_OK;
}

}  // extern "C"

// ------------------------------------------------- SpMV without allocation
// hprqp_hip_csr_spmv_noalloc keeps its row-block plan in `scratch`, a
// file-scope device array of fixed size (each GPU holds its own copy for the
// life of the process), instead of a per-call hipMalloc. Host code never
// passes `scratch` itself to the device: its address on each device comes
// from hipGetSymbolAddress on the first call there. The plan is built on the
// host from the row pointer read through a fixed window and uploaded in fixed
// chunks, so the call allocates neither device nor heap memory; a mutex
// serialises calls in one process (the plan and the staging are shared), and
// each call waits for its product before it returns, so a later call cannot
// overwrite a plan still in use.
constexpr int kNoAllocMaxRows = 1 << 19;  // rows per call; the plan needs <= rows + 1 entries
static __device__ __align__(256) int2 scratch[kNoAllocMaxRows + 1];

namespace {

constexpr int kNoAllocMaxDevices = 64;
constexpr int kNoAllocWindow = 8192;  // row-pointer entries read per D2H copy
constexpr int kNoAllocChunk = 4096;   // plan entries per H2D copy

struct NoAllocState {
    std::mutex lock;
    int2 *dev_scratch[kNoAllocMaxDevices] = {};
    int rp_window[kNoAllocWindow];
    int2 plan_chunk[kNoAllocChunk];
};
NoAllocState g_noalloc;

// Device address of `scratch` on the current device (resolved once per
// device, with that device current).
HPRQPStatus noalloc_scratch(int2 **out)
{
    int dev = 0;
    HPRQP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (!g_noalloc.dev_scratch[dev]) {
        void *ptr = nullptr;
        HPRQP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_noalloc.dev_scratch[dev] = static_cast<int2 *>(ptr);
    }
    *out = g_noalloc.dev_scratch[dev];
    return HPRQP_STATUS_OK;
}

}  // namespace

extern "C" {

HPRQPStatus hprqp_hip_csr_spmv_noalloc(int nrow, const int *row_ptr, const int *col_ind, const double *values,
                                       const double *x, double *y, void *stream)
{
    if (nrow < 0 || (nrow > 0 && (!row_ptr || !x || !y))) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (nrow == 0) return HPRQP_STATUS_OK;
    if (nrow > kNoAllocMaxRows) {
        std::fprintf(stderr, "hprqp_hip_csr_spmv_noalloc: %d rows exceed the limit of %d (use hprqp_hip_csr_spmv)\n",
                     nrow, kNoAllocMaxRows);
        return HPRQP_STATUS_INVALID_ARGUMENT;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> guard(g_noalloc.lock);
    int2 *plan = nullptr;
    HPRQPStatus st = noalloc_scratch(&plan);
    if (st != HPRQP_STATUS_OK) return st;
    // row pointer through a sliding window of kNoAllocWindow entries
    int base = 0, have = 0;
    hipError_t err = hipSuccess;
    auto rp = [&](int i) -> int {
        if (i < base || i >= base + have) {
            base = i > 0 ? i - 1 : 0;
            have = std::min(kNoAllocWindow, nrow + 1 - base);
            if (err == hipSuccess)
                err = hipMemcpyAsync(g_noalloc.rp_window, row_ptr + base, sizeof(int) * have, hipMemcpyDeviceToHost, s);
            if (err == hipSuccess) err = hipStreamSynchronize(s);
            if (err != hipSuccess) std::memset(g_noalloc.rp_window, 0, sizeof(int) * have);
        }
        return g_noalloc.rp_window[i - base];
    };
    int entries = 0, staged = 0;
    auto flush = [&]() {
        if (staged && err == hipSuccess)
            err = hipMemcpyAsync(plan + entries - staged, g_noalloc.plan_chunk, sizeof(int2) * staged,
                                 hipMemcpyHostToDevice, s);
        if (staged && err == hipSuccess) err = hipStreamSynchronize(s);
        staged = 0;
    };
    auto emit = [&](int2 e) {
        if (entries > kNoAllocMaxRows) return;  // cannot happen: one entry per block of >= 1 row, plus the end
        g_noalloc.plan_chunk[staged++] = e;
        ++entries;
        if (staged == kNoAllocChunk) flush();
    };
    hprqp::plan_row_blocks_with(nrow, rp, emit);
    flush();
    HPRQP_HIP(err);
    hprqp::EpiStore epi;
    epi.out = y;
    hprqp::spmv_rows_kernel<hprqp::EpiStore><<<entries - 1, hprqp::kThreads, 0, s>>>(row_ptr, col_ind, values, x,
                                                                                   plan, epi);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(hipStreamSynchronize(s));
    return HPRQP_STATUS_OK;
}

// A plain streaming copy of `bytes` (multiple of 16) between two fresh device
// buffers: reads and writes exactly `bytes` once. Reference kernel for
// calibrating the FETCH_SIZE / WRITE_SIZE counters.
__global__ void hprqp_calibration_copy_kernel(size_t n2, const double2 *__restrict__ src, double2 *__restrict__ dst)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

HPRQPStatus hprqp_hip_calibration_copy(size_t bytes, double *milliseconds)
{
    if (bytes == 0 || bytes % 16 != 0) return HPRQP_STATUS_INVALID_ARGUMENT;
    hprqp::DeviceBuffer<double2> a, b;
    HPRQP_HIP(a.alloc(bytes / 16));
    HPRQP_HIP(b.alloc(bytes / 16));
    HPRQP_HIP(hipMemset(a.p, 0, bytes));
    hprqp::EventPair ev;
    HPRQP_HIP(ev.create());
    HPRQP_HIP(hipEventRecord(ev.a, nullptr));
    hprqp_calibration_copy_kernel<<<2048, 256>>>(bytes / 16, a.p, b.p);
    HPRQP_HIP(hipEventRecord(ev.b, nullptr));
    HPRQP_HIP(hipEventSynchronize(ev.b));
    float ms = 0.0f;
    HPRQP_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
    if (milliseconds) *milliseconds = ms;
    return HPRQP_STATUS_OK;
}

// Random-gather reference kernel: an ELL product with `kGatherWidth` entries
// per row stored column-major (so index and value loads are coalesced) and
// columns spread uniformly over a window of the gathered vector. Its rate is
// the ceiling a gather-bound SpMV can reach for that window size.
constexpr int kGatherWidth = 20;

__global__ void hprqp_calibration_index_kernel(size_t count, unsigned window, int *__restrict__ idx,
                                               double *__restrict__ v)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x) {
        unsigned h = (unsigned)k * 2654435761u;
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        idx[k] = (int)(h % window);
        v[k] = 1.0;
    }
}

__global__ __launch_bounds__(256) void hprqp_calibration_gather_kernel(int nrow, const int *__restrict__ idx,
                                                                        const double *__restrict__ v,
                                                                        const double *__restrict__ x,
                                                                        double *__restrict__ y)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nrow) return;
    int c[kGatherWidth];
    double a[kGatherWidth];
#pragma unroll
    for (int j = 0; j < kGatherWidth; ++j) {
        c[j] = __builtin_nontemporal_load(idx + (size_t)j * nrow + r);
        a[j] = __builtin_nontemporal_load(v + (size_t)j * nrow + r);
    }
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kGatherWidth; ++j) acc = fma(a[j], x[c[j]], acc);
    y[r] = acc;
}

// Times one random-gather ELL product over `nrow` rows (20 entries each) whose
// columns fall in [0, window) of a vector of `window` doubles; reports the
// mean of `reps` launches in milliseconds.
HPRQPStatus hprqp_hip_calibration_gather(int nrow, int window, int reps, double *milliseconds)
{
    if (nrow <= 0 || window <= 0 || reps <= 0) return HPRQP_STATUS_INVALID_ARGUMENT;
    const size_t count = (size_t)nrow * kGatherWidth;
    hprqp::DeviceBuffer<int> idx;
    hprqp::DeviceBuffer<double> v, x, y;
    HPRQP_HIP(idx.alloc(count));
    HPRQP_HIP(v.alloc(count));
    HPRQP_HIP(x.alloc((size_t)window));
    HPRQP_HIP(y.alloc((size_t)nrow));
    HPRQP_HIP(hipMemset(x.p, 0, (size_t)window * sizeof(double)));
    hprqp_calibration_index_kernel<<<4096, 256>>>(count, (unsigned)window, idx.p, v.p);
    const int grid = (nrow + 255) / 256;
    hprqp_calibration_gather_kernel<<<grid, 256>>>(nrow, idx.p, v.p, x.p, y.p);  // warm-up
    hprqp::EventPair ev;
    HPRQP_HIP(ev.create());
    HPRQP_HIP(hipEventRecord(ev.a, nullptr));
    for (int k = 0; k < reps; ++k) hprqp_calibration_gather_kernel<<<grid, 256>>>(nrow, idx.p, v.p, x.p, y.p);
    HPRQP_HIP(hipEventRecord(ev.b, nullptr));
    HPRQP_HIP(hipEventSynchronize(ev.b));
    float ms = 0.0f;
    HPRQP_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
    if (milliseconds) *milliseconds = ms / reps;
    return HPRQP_STATUS_OK;
}

HPRQPStatus hprqp_hip_csr_transpose(int nrow, int ncol, int nnz, const int *row_ptr, const int *col_ind,
                                    const double *values, int *t_row_ptr, int *t_col_ind, double *t_values,
                                    int *perm, void *stream)
{
    if (nrow < 0 || ncol < 0 || nnz < 0 || !row_ptr || !t_row_ptr) return HPRQP_STATUS_INVALID_ARGUMENT;
    return hprqp::device_transpose(nrow, ncol, nnz, row_ptr, col_ind, values, t_row_ptr, t_col_ind, t_values, perm,
                                   static_cast<hipStream_t>(stream));
}

}  // extern "C"
