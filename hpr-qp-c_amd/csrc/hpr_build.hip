// build_from_QAbc's preprocessing on the GPU for large models
// (src/shared/build.c:157-252 symmetrisation, src/shared/sparse.c:81-125
// transpose), with the host path's results bit for bit.
//
// Symmetrisation: every stored entry (r, c, v) of Q emits (r, c, v/2) and then
// (c, r, v/2), in entry order; a stable radix sort by the key r * n + c gives
// the reference's (row, column) order with generation order among equal keys
// (its qsort over the same sequence, src/shared/build.c:168). Per key, the
// entries with |v| > 1e-15 are summed in that order by one thread, sums with
// |s| <= 1e-15 are dropped, and the survivors are compacted into CSR.
// Transpose: the device counting-sort transpose (rows ascending per column).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "device_build.h"
#include "device_common.h"
#include "hprqp.h"
#include "hprqp_hip.h"

namespace hprqp {
namespace {

using u64 = unsigned long long;
static_assert(sizeof(u64) == sizeof(size_t), "size_t indices are 64-bit");

constexpr size_t kMinDeviceNnz = (size_t)1 << 20;  // smaller models stay on the host
constexpr double kDropTol = 1e-15;                  // build.c:251

inline unsigned grid(size_t len) { return (unsigned)((len + kThreads - 1) / kThreads); }

bool device_build_applies(size_t nrow, size_t ncol, size_t nnz, size_t emitted)
{
    if (std::getenv("HPRQP_HOST_BUILD")) return false;
    if (nnz < kMinDeviceNnz || emitted > (size_t)INT_MAX || nrow > (size_t)INT_MAX || ncol > (size_t)INT_MAX)
        return false;
    const char *dev = hprqp_hip_device_name();
    return dev && dev[0] != '\0';
}

__global__ void sym_emit_kernel(int n, const u64 *__restrict__ rp, const u64 *__restrict__ ci,
                                const double *__restrict__ val, u64 *__restrict__ keys, double *__restrict__ vals,
                                int *__restrict__ bad)
{
    const int r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= n) return;
    const u64 k0 = rp[r], k1 = rp[r + 1];
    if (k1 < k0) {
        atomicOr(bad, 1);
        return;
    }
    for (u64 k = k0; k < k1; ++k) {
        u64 c = ci[k];
        if (c >= (u64)n) {
            atomicOr(bad, 1);
            c = 0;
        }
        const double v = 0.5 * val[k];
        keys[2 * k] = (u64)r * (u64)n + c;
        vals[2 * k] = v;
        keys[2 * k + 1] = c * (u64)n + (u64)r;
        vals[2 * k + 1] = v;
    }
}

// One thread per run of equal keys (its first entry): the in-order sum of
// the run's entries above the drop tolerance; keep = 1 when it survives.
__global__ void sym_merge_kernel(int N, const u64 *__restrict__ keys, const double *__restrict__ vals,
                                 double *__restrict__ sums, int *__restrict__ keep)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    int kp = 0;
    if (i == 0 || keys[i] != keys[i - 1]) {
        const u64 key = keys[i];
        double s = 0.0;
        bool any = false;
        for (int j = i; j < N && keys[j] == key; ++j) {
            const double v = vals[j];
            if (fabs(v) <= kDropTol) continue;
            s = any ? s + v : v;
            any = true;
        }
        if (any && !(fabs(s) <= kDropTol)) {
            kp = 1;
            sums[i] = s;
        }
    }
    keep[i] = kp;
}

__global__ void sym_scatter_kernel(int N, int n, const u64 *__restrict__ keys, const double *__restrict__ sums,
                                   const int *__restrict__ keep, const int *__restrict__ pos, u64 *__restrict__ col,
                                   double *__restrict__ val, int *__restrict__ row)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= N || !keep[i]) return;
    const int p = pos[i];
    col[p] = keys[i] % (u64)n;
    val[p] = sums[i];
    row[p] = (int)(keys[i] / (u64)n);
}

// Row pointers of M entries sorted by row: rp[r] = first entry of row >= r.
__global__ void row_ptr_kernel(int M, int nrow, const int *__restrict__ row, u64 *__restrict__ rp)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i > M) return;
    const int prev = i == 0 ? -1 : row[i - 1];
    const int cur = i == M ? nrow : row[i];
    for (int r = prev + 1; r <= cur; ++r) rp[r] = (u64)i;
}

__global__ void narrow_kernel(size_t len, const u64 *__restrict__ in, u64 bound, int *__restrict__ out,
                              int *__restrict__ bad)
{
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= len) return;
    const u64 v = in[i];
    if (v > bound) atomicOr(bad, 1);
    out[i] = (int)(v > bound ? 0 : v);
}

__global__ void widen_kernel(size_t len, const int *__restrict__ in, u64 *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < len) out[i] = (u64)in[i];
}

// HPRQP_BUILD_TRACE=1 reports each device step on stderr (diagnostics).
double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void trace(const char *what, size_t entries, double t0)
{
    if (std::getenv("HPRQP_BUILD_TRACE"))
        std::fprintf(stderr, "hprqp: %s on the device (%zu entries, %.4f s)\n", what, entries, now_s() - t0);
}

HPRQPStatus fail_host(HPRQPMatrixCSR *m, HPRQPStatus st)
{
    matrix_csr_free(m);
    return st;
}

}  // namespace

HPRQPStatus device_symmetrize_q(const HPRQPMatrixCSR *q, HPRQPMatrixCSR *out)
{
    const size_t n = q->nrow, nnz = q->row_ptr[n];
    if (!device_build_applies(n, n, nnz, 2 * nnz)) return HPRQP_STATUS_NOT_IMPLEMENTED;
    const double t0 = now_s();
    const int N = (int)(2 * nnz);
    DeviceBuffer<u64> rp, ci, keys, keys2, col;
    DeviceBuffer<double> val, vals, vals2, sums, oval;
    DeviceBuffer<int> bad, keep, pos, row;
    HPRQP_HIP(rp.alloc(n + 1));
    HPRQP_HIP(ci.alloc(nnz));
    HPRQP_HIP(val.alloc(nnz));
    HPRQP_HIP(keys.alloc(N));
    HPRQP_HIP(keys2.alloc(N));
    HPRQP_HIP(vals.alloc(N));
    HPRQP_HIP(vals2.alloc(N));
    HPRQP_HIP(bad.alloc(1));
    HPRQP_HIP(hipMemcpy(rp.p, q->row_ptr, sizeof(u64) * (n + 1), hipMemcpyHostToDevice));
    HPRQP_HIP(hipMemcpy(ci.p, q->col_ind, sizeof(u64) * nnz, hipMemcpyHostToDevice));
    HPRQP_HIP(hipMemcpy(val.p, q->values, sizeof(double) * nnz, hipMemcpyHostToDevice));
    HPRQP_HIP(hipMemset(bad.p, 0, sizeof(int)));
    sym_emit_kernel<<<grid(n), kThreads>>>((int)n, rp.p, ci.p, val.p, keys.p, vals.p, bad.p);
    HPRQP_HIP(hipGetLastError());
    int h_bad = 0;
    HPRQP_HIP(hipMemcpy(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost));
    if (h_bad) return HPRQP_STATUS_NOT_IMPLEMENTED;  // malformed pattern: the host path decides
    unsigned end_bit = 1;
    const u64 max_key = (u64)n * (u64)n - 1;
    while (end_bit < 64 && (max_key >> end_bit) != 0) ++end_bit;
    size_t tmp_bytes = 0;
    HPRQP_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys.p, keys2.p, vals.p, vals2.p, (size_t)N, 0, end_bit));
    DeviceBuffer<char> tmp;
    HPRQP_HIP(tmp.alloc(tmp_bytes));
    HPRQP_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, keys.p, keys2.p, vals.p, vals2.p, (size_t)N, 0, end_bit));
    HPRQP_HIP(sums.alloc(N));
    HPRQP_HIP(keep.alloc(N));
    HPRQP_HIP(pos.alloc(N));
    sym_merge_kernel<<<grid(N), kThreads>>>(N, keys2.p, vals2.p, sums.p, keep.p);
    HPRQP_HIP(hipGetLastError());
    size_t scan_bytes = 0;
    HPRQP_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, keep.p, pos.p, 0, (size_t)N, rocprim::plus<int>()));
    DeviceBuffer<char> scan_tmp;
    HPRQP_HIP(scan_tmp.alloc(scan_bytes));
    HPRQP_HIP(rocprim::exclusive_scan(scan_tmp.p, scan_bytes, keep.p, pos.p, 0, (size_t)N, rocprim::plus<int>()));
    int last_pos = 0, last_keep = 0;
    HPRQP_HIP(hipMemcpy(&last_pos, pos.p + (N - 1), sizeof(int), hipMemcpyDeviceToHost));
    HPRQP_HIP(hipMemcpy(&last_keep, keep.p + (N - 1), sizeof(int), hipMemcpyDeviceToHost));
    const int M = last_pos + last_keep;
    HPRQP_HIP(col.alloc(M));
    HPRQP_HIP(oval.alloc(M));
    HPRQP_HIP(row.alloc(M));
    sym_scatter_kernel<<<grid(N), kThreads>>>(N, (int)n, keys2.p, sums.p, keep.p, pos.p, col.p, oval.p, row.p);
    row_ptr_kernel<<<grid((size_t)M + 1), kThreads>>>(M, (int)n, row.p, rp.p);
    HPRQP_HIP(hipGetLastError());
    HPRQPStatus st = matrix_csr_alloc(out, n, q->ncol, (size_t)M);
    if (st != HPRQP_STATUS_OK) return st;
    const hipError_t e1 = hipMemcpy(out->row_ptr, rp.p, sizeof(u64) * (n + 1), hipMemcpyDeviceToHost);
    const hipError_t e2 = M ? hipMemcpy(out->col_ind, col.p, sizeof(u64) * M, hipMemcpyDeviceToHost) : hipSuccess;
    const hipError_t e3 = M ? hipMemcpy(out->values, oval.p, sizeof(double) * M, hipMemcpyDeviceToHost) : hipSuccess;
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail_host(out, HPRQP_STATUS_NOT_IMPLEMENTED);
    trace("Q symmetrised", nnz, t0);
    return HPRQP_STATUS_OK;
}

HPRQPStatus device_transpose_csr(const HPRQPMatrixCSR *a, HPRQPMatrixCSR *at)
{
    const size_t m = a->nrow, n = a->ncol, nnz = a->row_ptr[m];
    if (!device_build_applies(m, n, nnz, nnz)) return HPRQP_STATUS_NOT_IMPLEMENTED;
    const double t0 = now_s();
    DeviceBuffer<u64> w64;
    DeviceBuffer<int> rp, ci, trp, tci, bad;
    DeviceBuffer<double> val, tval;
    HPRQP_HIP(w64.alloc(std::max(nnz, std::max(m, n) + 1)));
    HPRQP_HIP(rp.alloc(m + 1));
    HPRQP_HIP(ci.alloc(nnz));
    HPRQP_HIP(val.alloc(nnz));
    HPRQP_HIP(trp.alloc(n + 1));
    HPRQP_HIP(tci.alloc(nnz));
    HPRQP_HIP(tval.alloc(nnz));
    HPRQP_HIP(bad.alloc(1));
    HPRQP_HIP(hipMemset(bad.p, 0, sizeof(int)));
    HPRQP_HIP(hipMemcpy(w64.p, a->row_ptr, sizeof(u64) * (m + 1), hipMemcpyHostToDevice));
    narrow_kernel<<<grid(m + 1), kThreads>>>(m + 1, w64.p, (u64)nnz, rp.p, bad.p);
    HPRQP_HIP(hipMemcpy(w64.p, a->col_ind, sizeof(u64) * nnz, hipMemcpyHostToDevice));
    narrow_kernel<<<grid(nnz), kThreads>>>(nnz, w64.p, (u64)(n - 1), ci.p, bad.p);
    HPRQP_HIP(hipMemcpy(val.p, a->values, sizeof(double) * nnz, hipMemcpyHostToDevice));
    HPRQP_HIP(hipGetLastError());
    int h_bad = 0;
    HPRQP_HIP(hipMemcpy(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost));
    if (h_bad || n == 0) return HPRQP_STATUS_NOT_IMPLEMENTED;
    HPRQPStatus st = hprqp_hip_csr_transpose((int)m, (int)n, (int)nnz, rp.p, ci.p, val.p, trp.p, tci.p, tval.p,
                                             nullptr, nullptr);
    if (st != HPRQP_STATUS_OK) return st;
    st = matrix_csr_alloc(at, n, m, nnz);
    if (st != HPRQP_STATUS_OK) return st;
    widen_kernel<<<grid(n + 1), kThreads>>>(n + 1, trp.p, w64.p);
    if (hipMemcpy(at->row_ptr, w64.p, sizeof(u64) * (n + 1), hipMemcpyDeviceToHost) != hipSuccess)
        return fail_host(at, HPRQP_STATUS_NOT_IMPLEMENTED);
    widen_kernel<<<grid(nnz), kThreads>>>(nnz, tci.p, w64.p);
    if (hipMemcpy(at->col_ind, w64.p, sizeof(u64) * nnz, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(at->values, tval.p, sizeof(double) * nnz, hipMemcpyDeviceToHost) != hipSuccess)
        return fail_host(at, HPRQP_STATUS_NOT_IMPLEMENTED);
    trace("A transposed", nnz, t0);
    return HPRQP_STATUS_OK;
}

}  // namespace hprqp
