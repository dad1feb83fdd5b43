// Equality-row (Gram) metric on gfx950: host setup (metric rows, beta, the
// assembly of G = E E', validation) and the device factorisation and solves.
// The formulas follow src/gpu/qp_row_metric.cu:94-149 (plain QPs),
// src/gpu/row_metric.cu:68-130 (row-cone prefix) and
// src/shared/gram_policy.cpp:6-50 (AUTO); the linear algebra is this file's
// own: a blocked right-looking Cholesky on the device, the explicit
// triangular inverse W = L^{-1}, and per-step solves x = W' (W r) as two
// HBM-streaming triangular matrix-vector products.
#include "gram.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "device_common.h"

namespace hprqp {

namespace {

constexpr int kNB = 32;    // Cholesky / inverse block
constexpr int kTile = 64;  // output tile of the block updates (4 x 4 per thread)

bool diagnostics_on()
{
    const char *v = std::getenv("HPRQP_DIAGNOSTICS");
    return v && std::strcmp(v, "1") == 0;
}

// ------------------------------------------------------------ factorisation

// Cholesky of the diagonal block [k0, k0 + kb)^2 in place (lower part). A
// pivot that is not positive and finite records its 1-based index in *info
// (first one wins) and is replaced by 1 so that the sweep stays finite.
__global__ __launch_bounds__(kThreads) void gram_potrf_diag_kernel(double *A, int ld, int k0, int kb, int *info)
{
    __shared__ double s[kNB][kNB + 1];
    const int t = threadIdx.x;
    for (int idx = t; idx < kb * kb; idx += kThreads) {
        const int i = idx / kb, j = idx % kb;
        s[i][j] = A[(size_t)(k0 + i) * ld + k0 + j];
    }
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
        if (t == 0) {
            double d = s[j][j];
            if (!(d > 0.0) || !isfinite(d)) {
                if (*info == 0) *info = k0 + j + 1;
                d = 1.0;
            }
            s[j][j] = sqrt(d);
        }
        __syncthreads();
        for (int i = j + 1 + t; i < kb; i += kThreads) s[i][j] /= s[j][j];
        __syncthreads();
        const int rem = kb - j - 1;
        for (int idx = t; idx < rem * rem; idx += kThreads) {
            const int i = j + 1 + idx / rem, c = j + 1 + idx % rem;
            if (c <= i) s[i][c] -= s[i][j] * s[c][j];
        }
        __syncthreads();
    }
    for (int idx = t; idx < kb * kb; idx += kThreads) {
        const int i = idx / kb, j = idx % kb;
        A[(size_t)(k0 + i) * ld + k0 + j] = j <= i ? s[i][j] : 0.0;
    }
}

// Loads the lower part of the factor's diagonal block [k0, k0 + kb)^2.
__device__ void load_diag_block(double (*s)[kNB + 1], const double *L, int ld, int k0, int kb)
{
    for (int idx = threadIdx.x; idx < kNB * kNB; idx += kThreads) {
        const int i = idx / kNB, j = idx % kNB;
        s[i][j] = (i < kb && j <= i) ? L[(size_t)(k0 + i) * ld + k0 + j] : 0.0;
    }
}

// Panel below the diagonal block: row i of L[:, k0:k0+kb] solves
// x L_kk' = a (forward over the block's columns), one row per thread.
__global__ __launch_bounds__(kThreads) void gram_trsm_panel_kernel(double *A, int ld, int e, int k0, int kb)
{
    __shared__ double s[kNB][kNB + 1];
    load_diag_block(s, A, ld, k0, kb);
    __syncthreads();
    const int i = k0 + kb + blockIdx.x * kThreads + threadIdx.x;
    if (i >= e) return;
    double *row = A + (size_t)i * ld + k0;
    double x[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c) x[c] = c < kb ? row[c] : 0.0;
#pragma unroll
    for (int c = 0; c < kNB; ++c) {
        if (c < kb) {
            double v = x[c];
#pragma unroll
            for (int q = 0; q < c; ++q) v -= x[q] * s[c][q];
            x[c] = v / s[c][c];
        }
    }
#pragma unroll
    for (int c = 0; c < kNB; ++c)
        if (c < kb) row[c] = x[c];
}

// Trailing update A[i][j] -= sum_t L[i][k0+t] L[j][k0+t] on the lower
// triangle j <= i of [k0+kb, e)^2, one 64 x 64 tile per workgroup (lower
// tiles only, numbered row by row), 4 x 4 outputs per thread.
__global__ __launch_bounds__(kThreads) void gram_syrk_kernel(double *A, int ld, int e, int k0, int kb)
{
    __shared__ double si[kTile][kNB + 1], sj[kTile][kNB + 1];
    const int j0 = k0 + kb;
    const int b = blockIdx.x;
    int ti = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
    while (ti * (ti + 1) / 2 > b) --ti;
    const int tj = b - ti * (ti + 1) / 2;
    const int r0 = j0 + ti * kTile, c0 = j0 + tj * kTile;
    for (int idx = threadIdx.x; idx < kTile * kNB; idx += kThreads) {
        const int r = idx / kNB, t = idx % kNB;
        si[r][t] = (r0 + r < e && t < kb) ? A[(size_t)(r0 + r) * ld + k0 + t] : 0.0;
        sj[r][t] = (c0 + r < e && t < kb) ? A[(size_t)(c0 + r) * ld + k0 + t] : 0.0;
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4] = {};
    for (int t = 0; t < kb; ++t) {
        double a[4], bb[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            a[p] = si[ty * 4 + p][t];
            bb[p] = sj[tx * 4 + p][t];
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * bb[q];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = r0 + ty * 4 + p, j = c0 + tx * 4 + q;
            if (i < e && j <= i) A[(size_t)i * ld + j] -= acc[p][q];
        }
}

// W := I on the ld x ld buffer (rows and columns past e stay zero).
__global__ __launch_bounds__(kThreads) void gram_identity_kernel(double *W, int ld, int e)
{
    const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (idx >= (size_t)ld * ld) return;
    const int i = (int)(idx / ld), j = (int)(idx % ld);
    W[idx] = (i == j && i < e) ? 1.0 : 0.0;
}

// Block row k of X = L^{-1} B (B = I updated in place): X_k = L_kk^{-1} B_k
// over columns [0, ncol), one column per thread (forward substitution).
__global__ __launch_bounds__(kThreads) void gram_trsm_left_kernel(const double *L, double *W, int ld, int k0, int kb,
                                                                  int ncol)
{
    __shared__ double s[kNB][kNB + 1];
    load_diag_block(s, L, ld, k0, kb);
    __syncthreads();
    const int j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= ncol) return;
    double x[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c) x[c] = c < kb ? W[(size_t)(k0 + c) * ld + j] : 0.0;
#pragma unroll
    for (int c = 0; c < kNB; ++c) {
        if (c < kb) {
            double v = x[c];
#pragma unroll
            for (int q = 0; q < c; ++q) v -= s[c][q] * x[q];
            x[c] = v / s[c][c];
        }
    }
#pragma unroll
    for (int c = 0; c < kNB; ++c)
        if (c < kb) W[(size_t)(k0 + c) * ld + j] = x[c];
}

// Rows below block k: B_i[:, 0:ncol] -= L[i][k0:k0+kb] X_k[:, 0:ncol] for
// i in [k0+kb, e), 64 x 64 tiles (tile index row-major over the tile grid).
__global__ __launch_bounds__(kThreads) void gram_inv_update_kernel(const double *L, double *W, int ld, int e, int k0,
                                                                   int kb, int ncol, int tiles_c)
{
    __shared__ double sa[kTile][kNB + 1];
    __shared__ double sb[kNB][kTile + 1];
    const int tr = blockIdx.x / tiles_c, tc = blockIdx.x % tiles_c;
    const int r0 = k0 + kb + tr * kTile, c0 = tc * kTile;
    for (int idx = threadIdx.x; idx < kTile * kNB; idx += kThreads) {
        const int r = idx / kNB, t = idx % kNB;
        sa[r][t] = (r0 + r < e && t < kb) ? L[(size_t)(r0 + r) * ld + k0 + t] : 0.0;
        const int t2 = idx / kTile, c = idx % kTile;
        sb[t2][c] = (t2 < kb && c0 + c < ncol) ? W[(size_t)(k0 + t2) * ld + c0 + c] : 0.0;
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4] = {};
    for (int t = 0; t < kb; ++t) {
        double a[4], bb[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            a[p] = sa[ty * 4 + p][t];
            bb[p] = sb[t][tx * 4 + p];
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * bb[q];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = r0 + ty * 4 + p, j = c0 + tx * 4 + q;
            if (i < e && j < ncol) W[(size_t)i * ld + j] -= acc[p][q];
        }
}

// WT = W' on ld x ld (32 x 32 tiles through LDS).
__global__ __launch_bounds__(kThreads) void gram_transpose_kernel(const double *W, double *WT, int ld)
{
    __shared__ double s[32][33];
    const int tiles = ld / 32;
    const int bi = blockIdx.x / tiles, bj = blockIdx.x % tiles;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) s[r][tx] = W[(size_t)(bi * 32 + r) * ld + bj * 32 + tx];
    __syncthreads();
    for (int r = ty; r < 32; r += 8) WT[(size_t)(bj * 32 + r) * ld + bi * 32 + tx] = s[tx][r];
}

// ------------------------------------------------------------------ solves

// u[q] = sum_{k <= q} W[q][k] r[k]: one wavefront per row, four chains per
// lane, a DPP wavefront sum (every lane of a live wavefront is active).
__global__ __launch_bounds__(kThreads) void gram_lower_gemv_kernel(int e, int ld, const double *__restrict__ W,
                                                                   const double *__restrict__ r,
                                                                   double *__restrict__ u)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * (kThreads / kWave) + (threadIdx.x >> 6);
    if (q >= e) return;
    const double *row = W + (size_t)q * ld;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int k = lane;
    for (; k + 3 * kWave <= q; k += 4 * kWave) {
        a0 += row[k] * r[k];
        a1 += row[k + kWave] * r[k + kWave];
        a2 += row[k + 2 * kWave] * r[k + 2 * kWave];
        a3 += row[k + 3 * kWave] * r[k + 3 * kWave];
    }
    for (; k <= q; k += kWave) a0 += row[k] * r[k];
    const double v = wave_sum_dpp((a0 + a1) + (a2 + a3));
    if (lane == 0) u[q] = v;
}

// step[j] = sum_{k >= j} WT[j][k] u[k] (= (W' u)_j); with `finish` set the
// metric row rows[j] takes its y update (qp_row_metric.cu:55-60).
__global__ __launch_bounds__(kThreads) void gram_upper_gemv_kernel(int e, int ld, const double *__restrict__ WT,
                                                                   const double *__restrict__ u,
                                                                   double *__restrict__ step, int finish,
                                                                   const int *__restrict__ rows, double *y,
                                                                   double *y_bar, double *dy,
                                                                   const double *__restrict__ last_y, double sigma,
                                                                   double a, double b)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int j = blockIdx.x * (kThreads / kWave) + (threadIdx.x >> 6);
    if (j >= e) return;
    const double *row = WT + (size_t)j * ld;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int k = j + lane;
    for (; k + 3 * kWave < e; k += 4 * kWave) {
        a0 += row[k] * u[k];
        a1 += row[k + kWave] * u[k + kWave];
        a2 += row[k + 2 * kWave] * u[k + 2 * kWave];
        a3 += row[k + 3 * kWave] * u[k + 3 * kWave];
    }
    for (; k < e; k += kWave) a0 += row[k] * u[k];
    const double v = wave_sum_dpp((a0 + a1) + (a2 + a3));
    if (lane == 0) {
        step[j] = v;
        if (finish) {
            const int i = rows[j];
            const double yi = y[i];
            const double yb = yi - v / (2.0 * sigma);
            dy[i] = yb - yi;
            y_bar[i] = yb;
            y[i] = a * last_y[i] + b * (2.0 * yb - yi);
        }
    }
}

// Per-workgroup partial sums of the quadratic: slot 0 of block g holds
// sum (E' dy)_j^2 over its rows j of E', slot 1 sum dy_i^2 over the rows i
// outside the metric in its range of rows of A.
__global__ __launch_bounds__(kThreads) void gram_quadratic_kernel(int n, int m, const int *__restrict__ rp,
                                                                  const int *__restrict__ ci,
                                                                  const double *__restrict__ val,
                                                                  const int *__restrict__ map,
                                                                  const double *__restrict__ dy, double *part)
{
    const int t = blockIdx.x * kThreads + threadIdx.x;
    double eq = 0.0, rest = 0.0;
    if (t < n) {
        double v = 0.0;
        for (int k = rp[t]; k < rp[t + 1]; ++k) v += val[k] * dy[ci[k]];
        eq = v * v;
    }
    if (t < m && map[t] < 0) rest = dy[t] * dy[t];
    eq = block_sum_all(eq);
    rest = block_sum_all(rest);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = eq;
        part[2 * blockIdx.x + 1] = rest;
    }
}

}  // namespace

// ------------------------------------------------------------------ host side

GramMetric::~GramMetric() { release(); }

void GramMetric::release()
{
    for (double *p : {d_W_, d_WT_, d_r_, d_u_, d_step_, d_b_, d_et_val_, d_part_})
        if (p) (void)hipFree(p);
    for (int *p : {d_rows_, d_map_, d_et_rp_, d_et_ci_})
        if (p) (void)hipFree(p);
    d_W_ = d_WT_ = d_r_ = d_u_ = d_step_ = d_b_ = d_et_val_ = d_part_ = nullptr;
    d_rows_ = d_map_ = d_et_rp_ = d_et_ci_ = nullptr;
    kind_ = e_ = 0;
}

// G (ld_ x ld_, dense, lower part used) -> W = L^{-1} and W' on the device.
HPRQPStatus GramMetric::factor(const std::vector<double> &G, hipStream_t s, bool *spd)
{
    const int e = e_, ld = ld_;
    const size_t cells = (size_t)ld * ld;
    DeviceBuffer<double> L;
    DeviceBuffer<int> info;
    HPRQP_HIP(L.alloc(cells));
    HPRQP_HIP(info.alloc(1));
    HPRQP_HIP(hipMemcpyAsync(L.p, G.data(), sizeof(double) * cells, hipMemcpyHostToDevice, s));
    HPRQP_HIP(hipMemsetAsync(info.p, 0, sizeof(int), s));
    for (int k0 = 0; k0 < e; k0 += kNB) {
        const int kb = std::min(kNB, e - k0);
        gram_potrf_diag_kernel<<<1, kThreads, 0, s>>>(L.p, ld, k0, kb, info.p);
        const int rest = e - k0 - kb;
        if (rest > 0) {
            gram_trsm_panel_kernel<<<(rest + kThreads - 1) / kThreads, kThreads, 0, s>>>(L.p, ld, e, k0, kb);
            const int T = (rest + kTile - 1) / kTile;
            gram_syrk_kernel<<<T * (T + 1) / 2, kThreads, 0, s>>>(L.p, ld, e, k0, kb);
        }
    }
    HPRQP_HIP(hipGetLastError());
    int h_info = 0;
    HPRQP_HIP(hipMemcpyAsync(&h_info, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HPRQP_HIP(hipStreamSynchronize(s));
    *spd = h_info == 0;
    if (!*spd) return HPRQP_STATUS_OK;
    HPRQP_HIP(hipMalloc((void **)&d_W_, sizeof(double) * cells));
    HPRQP_HIP(hipMalloc((void **)&d_WT_, sizeof(double) * cells));
    gram_identity_kernel<<<(unsigned)((cells + kThreads - 1) / kThreads), kThreads, 0, s>>>(d_W_, ld, e);
    for (int k0 = 0; k0 < e; k0 += kNB) {
        const int kb = std::min(kNB, e - k0);
        const int ncol = k0 + kb;
        gram_trsm_left_kernel<<<(ncol + kThreads - 1) / kThreads, kThreads, 0, s>>>(L.p, d_W_, ld, k0, kb, ncol);
        const int rest = e - k0 - kb;
        if (rest > 0) {
            const int tiles_c = (ncol + kTile - 1) / kTile, tiles_r = (rest + kTile - 1) / kTile;
            gram_inv_update_kernel<<<tiles_r * tiles_c, kThreads, 0, s>>>(L.p, d_W_, ld, e, k0, kb, ncol, tiles_c);
        }
    }
    const int tiles = ld / 32;
    gram_transpose_kernel<<<tiles * tiles, kThreads, 0, s>>>(d_W_, d_WT_, ld);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(hipStreamSynchronize(s));
    return HPRQP_STATUS_OK;
}

// x = G^{-1} rhs on the device (setup-time checks).
HPRQPStatus GramMetric::solve(const std::vector<double> &rhs, std::vector<double> &x, hipStream_t s)
{
    const int grid = (e_ + kThreads / kWave - 1) / (kThreads / kWave);
    HPRQP_HIP(hipMemcpyAsync(d_r_, rhs.data(), sizeof(double) * e_, hipMemcpyHostToDevice, s));
    gram_lower_gemv_kernel<<<grid, kThreads, 0, s>>>(e_, ld_, d_W_, d_r_, d_u_);
    gram_upper_gemv_kernel<<<grid, kThreads, 0, s>>>(e_, ld_, d_WT_, d_u_, d_step_, 0, nullptr, nullptr, nullptr,
                                                     nullptr, nullptr, 1.0, 0.0, 0.0);
    HPRQP_HIP(hipGetLastError());
    x.resize(e_);
    HPRQP_HIP(hipMemcpyAsync(x.data(), d_step_, sizeof(double) * e_, hipMemcpyDeviceToHost, s));
    HPRQP_HIP(hipStreamSynchronize(s));
    return HPRQP_STATUS_OK;
}

HPRQPStatus GramMetric::setup(const GramInputs &in, hipStream_t s, double *sigma, double *lambda_A,
                              const char **reason)
{
    release();
    *reason = nullptr;
    const int n = in.n, m = in.m;
    n_ = n;
    m_ = m;
    int kind = 0;
    std::vector<int> map(m, -1), rows;
    std::vector<double> b;
    long double rest_bound = 0.0L;
    double beta = 0.0;
    std::vector<double> cols(n, 0.0);
    double rowmax = 0.0;
    if (!in.num_var_cones && !in.num_row_cones) {
        // qp_row_metric.cu:94-108: the rows with AL == AU finite
        if (!m || !n || in.rp[m] == 0) {
            *reason = "no_equalities";
            return HPRQP_STATUS_OK;
        }
        for (int i = 0; i < m; ++i)
            if (std::isfinite(in.AL[i]) && in.AL[i] == in.AU[i]) {
                map[i] = (int)rows.size();
                rows.push_back(i);
                b.push_back(in.AL[i]);
            }
        if (rows.empty()) {
            *reason = "no_equalities";
            return HPRQP_STATUS_OK;
        }
        if ((int)rows.size() > n) {
            *reason = "more_equalities_than_variables";
            return HPRQP_STATUS_OK;
        }
        for (int i = 0; i < m; ++i) {
            if (map[i] >= 0) continue;
            double row = 0.0;
            for (int k = in.rp[i]; k < in.rp[i + 1]; ++k) {
                row += std::fabs(in.val[k]);
                cols[in.ci[k]] += std::fabs(in.val[k]);
            }
            rowmax = std::max(rowmax, row);
            const double lo = std::isfinite(in.AL[i]) ? std::fabs(in.AL[i]) : 0.0;
            const double hi = std::isfinite(in.AU[i]) ? std::fabs(in.AU[i]) : 0.0;
            const double v = std::max(lo, hi);
            rest_bound += (long double)v * v;
        }
        beta = (1 + 1e-12) * rowmax * (n ? *std::max_element(cols.begin(), cols.end()) : 0.0);
        if (beta == 0.0) beta = 1.0;
        if (!std::isfinite(beta)) {
            *reason = "nonfinite_majorant";
            return HPRQP_STATUS_OK;
        }
        kind = 1;
    } else if (in.num_row_cones) {
        // row_metric.cu:68-80: a homogeneous equality prefix, diagonal Q
        const int l = in.lin_m;
        if (in.num_var_cones || !l || !in.Q_is_diag || in.noC) {
            *reason = "soc_setup_rejected";
            return HPRQP_STATUS_OK;
        }
        for (int i = 0; i < l; ++i)
            if (in.AL[i] != 0.0 || in.AU[i] != 0.0 || (in.soc_rhs && in.soc_rhs[i] != 0.0)) {
                *reason = "soc_setup_rejected";
                return HPRQP_STATUS_OK;
            }
        for (int i = l; i < m; ++i) {
            double row = 0.0;
            for (int k = in.rp[i]; k < in.rp[i + 1]; ++k) {
                row += std::fabs(in.val[k]);
                cols[in.ci[k]] += std::fabs(in.val[k]);
            }
            rowmax = std::max(rowmax, row);
        }
        beta = (1 + 1e-12) * rowmax * (n ? *std::max_element(cols.begin(), cols.end()) : 0.0);
        if (!(beta > 0.0)) {
            *reason = "nonfinite_majorant";
            return HPRQP_STATUS_OK;
        }
        for (int i = 0; i < l; ++i) {
            map[i] = i;
            rows.push_back(i);
            b.push_back(0.0);
        }
        kind = 2;
    } else {
        *reason = "variable_cones";
        return HPRQP_STATUS_OK;
    }
    const int e = (int)rows.size();
    if (e > kGramMaxRows) {
        *reason = "gram_rows_limit";
        return HPRQP_STATUS_OK;
    }
    // E' by a counting sort over the metric rows (column = row of A), and the
    // assembly-work limit of qp_row_metric.cu:102-103
    std::vector<int> et_rp(n + 1, 0);
    for (int i : rows)
        for (int k = in.rp[i]; k < in.rp[i + 1]; ++k) et_rp[in.ci[k] + 1] += 1;
    unsigned long long work = 0;
    for (int j = 0; j < n; ++j) work += (unsigned long long)et_rp[j + 1] * et_rp[j + 1];
    if (work > 100000000ULL) {
        *reason = "gram_assembly_work_limit";
        return HPRQP_STATUS_OK;
    }
    for (int j = 0; j < n; ++j) et_rp[j + 1] += et_rp[j];
    std::vector<int> et_ci(et_rp[n]), cursor(et_rp.begin(), et_rp.end() - 1);
    std::vector<double> et_val(et_rp[n]);
    for (int i : rows)
        for (int k = in.rp[i]; k < in.rp[i + 1]; ++k) {
            const int z = cursor[in.ci[k]]++;
            et_ci[z] = i;
            et_val[z] = in.val[k];
        }
    // G row by row: the products of row i with every metric row sharing a
    // column, summed per column in the order they are generated
    const int ld = std::max(32, (e + 31) / 32 * 32);
    std::vector<double> G((size_t)ld * ld, 0.0);
    std::vector<int> g_rp(e + 1, 0), g_ci;
    std::vector<double> g_val;
    {
        std::vector<std::pair<int, double>> pairs;
        for (int q = 0; q < e; ++q) {
            const int i = rows[q];
            pairs.clear();
            for (int k = in.rp[i]; k < in.rp[i + 1]; ++k) {
                const int j = in.ci[k];
                for (int z = et_rp[j]; z < et_rp[j + 1]; ++z) pairs.emplace_back(map[et_ci[z]], in.val[k] * et_val[z]);
            }
            std::stable_sort(pairs.begin(), pairs.end(),
                             [](const std::pair<int, double> &x, const std::pair<int, double> &y) { return x.first < y.first; });
            for (size_t z = 0; z < pairs.size();) {
                const int col = pairs[z].first;
                double sum = 0.0;
                do {
                    sum += pairs[z++].second;
                } while (z < pairs.size() && pairs[z].first == col);
                if (sum != 0.0) {
                    G[(size_t)q * ld + col] = sum;
                    g_ci.push_back(col);
                    g_val.push_back(sum);
                }
            }
            g_rp[q + 1] = (int)g_ci.size();
        }
    }
    e_ = e;
    ld_ = ld;
    beta_ = beta;
    HPRQP_HIP(hipMalloc((void **)&d_r_, sizeof(double) * ld));
    HPRQP_HIP(hipMalloc((void **)&d_u_, sizeof(double) * ld));
    HPRQP_HIP(hipMalloc((void **)&d_step_, sizeof(double) * ld));
    bool spd = false;
    HPRQPStatus st = factor(G, s, &spd);
    if (st != HPRQP_STATUS_OK) return st;
    if (!spd) {
        release();
        *reason = "gram_not_spd";
        return HPRQP_STATUS_OK;
    }
    // validation solve, qp_row_metric.cu:120-127 / row_metric.cu:98-103
    std::vector<double> t(e), x;
    for (int i = 0; i < e; ++i) t[i] = std::sin(.173 * (i + 1)) + .2 * std::cos(.719 * (i + 1));
    st = solve(t, x, s);
    if (st != HPRQP_STATUS_OK) return st;
    double err = 0.0, an = 0.0, xn = 0.0, bn = 0.0;
    for (int i = 0; i < e; ++i) {
        if (!std::isfinite(x[i])) {
            release();
            *reason = "nonfinite_test_solve";
            return HPRQP_STATUS_OK;
        }
        long double z = 0.0L;
        double rs = 0.0;
        for (int k = g_rp[i]; k < g_rp[i + 1]; ++k) {
            z += (long double)g_val[k] * x[g_ci[k]];
            rs += std::fabs(g_val[k]);
        }
        err = std::max(err, (double)std::fabs(z - (long double)t[i]));
        an = std::max(an, rs);
        xn = std::max(xn, std::fabs(x[i]));
        bn = std::max(bn, std::fabs(t[i]));
    }
    const double backward = err / (an * xn + bn), relative = err / bn;
    if (!(backward <= 1e-10) || (e <= 5000 && !(relative <= 1e-8))) {
        release();
        *reason = "inverse_validation";
        return HPRQP_STATUS_OK;
    }
    double new_sigma = *sigma;
    if (kind == 1) {
        // qp_row_metric.cu:128-133: sigma from the metric norm of b
        st = solve(b, x, s);
        if (st != HPRQP_STATUS_OK) return st;
        long double bq = 0.0L;
        for (int i = 0; i < e; ++i) bq += (long double)b[i] * x[i];
        if (!std::isfinite((double)bq) || bq < -1e-10L) {
            release();
            *reason = "rhs_metric_norm";
            return HPRQP_STATUS_OK;
        }
        const double normb = std::sqrt((double)((bq > 0.0L ? bq : 0.0L) + rest_bound / beta));
        double c2 = 0.0;
        for (int j = 0; j < n; ++j) c2 += in.c[j] * in.c[j];
        const double normc = std::sqrt(c2);
        new_sigma = (normb > 1e-16 && normc > 1e-16 && normb < 1e16 && normc < 1e16) ? normb / normc : 1.0;
    } else {
        new_sigma = *sigma / std::sqrt(beta);  // row_metric.cu:128
    }
    HPRQP_HIP(hipMalloc((void **)&d_rows_, sizeof(int) * e));
    HPRQP_HIP(hipMalloc((void **)&d_b_, sizeof(double) * e));
    HPRQP_HIP(hipMalloc((void **)&d_map_, sizeof(int) * (m ? m : 1)));
    HPRQP_HIP(hipMalloc((void **)&d_et_rp_, sizeof(int) * (n + 1)));
    HPRQP_HIP(hipMalloc((void **)&d_et_ci_, sizeof(int) * std::max(1, et_rp[n])));
    HPRQP_HIP(hipMalloc((void **)&d_et_val_, sizeof(double) * std::max(1, et_rp[n])));
    part_blocks_ = (std::max(n, m) + kThreads - 1) / kThreads;
    HPRQP_HIP(hipMalloc((void **)&d_part_, sizeof(double) * 2 * std::max(1, part_blocks_)));
    part_h_.assign(2 * std::max(1, part_blocks_), 0.0);
    HPRQP_HIP(hipMemcpyAsync(d_rows_, rows.data(), sizeof(int) * e, hipMemcpyHostToDevice, s));
    HPRQP_HIP(hipMemcpyAsync(d_b_, b.data(), sizeof(double) * e, hipMemcpyHostToDevice, s));
    if (m) HPRQP_HIP(hipMemcpyAsync(d_map_, map.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
    HPRQP_HIP(hipMemcpyAsync(d_et_rp_, et_rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (et_rp[n]) {
        HPRQP_HIP(hipMemcpyAsync(d_et_ci_, et_ci.data(), sizeof(int) * et_rp[n], hipMemcpyHostToDevice, s));
        HPRQP_HIP(hipMemcpyAsync(d_et_val_, et_val.data(), sizeof(double) * et_rp[n], hipMemcpyHostToDevice, s));
    }
    HPRQP_HIP(hipStreamSynchronize(s));
    kind_ = kind;
    *sigma = new_sigma;
    *lambda_A = 2.0 * beta;
    if (diagnostics_on())
        std::fprintf(stderr,
                     "GRAM_READY kind=%s equality_rows=%d n=%d gram_nnz=%zu beta=%.17g sigma=%.17g backward=%.3g "
                     "rhs_relative=%.3g\n",
                     kind == 1 ? "qp" : "row_cone", e, n, g_val.size(), beta, new_sigma, backward, relative);
    return HPRQP_STATUS_OK;
}

HPRQPStatus GramMetric::finish_y(hipStream_t s, double *y, double *y_bar, double *dy, const double *last_y,
                                 double sigma, double a, double b)
{
    const int grid = (e_ + kThreads / kWave - 1) / (kThreads / kWave);
    gram_lower_gemv_kernel<<<grid, kThreads, 0, s>>>(e_, ld_, d_W_, d_r_, d_u_);
    gram_upper_gemv_kernel<<<grid, kThreads, 0, s>>>(e_, ld_, d_WT_, d_u_, d_step_, 1, d_rows_, y, y_bar, dy, last_y,
                                                     sigma, a, b);
    HPRQP_HIP(hipGetLastError());
    return HPRQP_STATUS_OK;
}

HPRQPStatus GramMetric::quadratic(hipStream_t s, const double *dy, double *out)
{
    gram_quadratic_kernel<<<std::max(1, part_blocks_), kThreads, 0, s>>>(n_, m_, d_et_rp_, d_et_ci_, d_et_val_,
                                                                         d_map_, dy, d_part_);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(hipMemcpyAsync(part_h_.data(), d_part_, sizeof(double) * part_h_.size(), hipMemcpyDeviceToHost, s));
    HPRQP_HIP(hipStreamSynchronize(s));
    double eq = 0.0, rest = 0.0;
    for (int g = 0; g < std::max(1, part_blocks_); ++g) {
        eq += part_h_[2 * g];
        rest += part_h_[2 * g + 1];
    }
    *out = 2.0 * eq + 2.0 * beta_ * rest;
    return std::isfinite(*out) ? HPRQP_STATUS_OK : HPRQP_STATUS_INVALID_ARGUMENT;
}

size_t GramMetric::read(const char *name, double *out, size_t cap, hipStream_t s) const
{
    const double *src = nullptr;
    if (!kind_) return 0;
    if (!std::strcmp(name, "gram_r")) src = d_r_;
    else if (!std::strcmp(name, "gram_step")) src = d_step_;
    else if (!std::strcmp(name, "gram_b")) src = d_b_;
    if (!src) return 0;
    if (out && cap) {
        const size_t cnt = std::min(cap, (size_t)e_);
        if (hipMemcpyAsync(out, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return 0;
    }
    return (size_t)e_;
}

// gram_policy.cpp:6-50
bool gram_auto_eligible(const HPRQPModel *md)
{
    if (!md || md->num_var_cones || !md->num_row_cones || !md->row_cone_offsets) return false;
    const size_t e = md->row_cone_offsets[0], n = md->A.ncol;
    const size_t g = (size_t)std::sqrt((double)e);
    if (g < 2 || g * g != e || e > n) return false;
    for (size_t k = 0; k < md->num_row_cones; ++k)
        if (md->row_cone_offsets[k + 1] - md->row_cone_offsets[k] != 3) return false;
    for (size_t i = 0; i < e; ++i)
        if (md->AL[i] != 0.0 || md->AU[i] != 0.0 || (md->soc_rhs && md->soc_rhs[i] != 0.0)) return false;
    const size_t *rp = md->A.row_ptr, *ci = md->A.col_ind;
    const double *av = md->A.values;
    std::vector<size_t> degree(n, 0), location(n, e), state(e, n);
    for (size_t k = 0; k < rp[e]; ++k) degree[ci[k]] += 1;
    if ((size_t)std::count_if(degree.begin(), degree.end(), [](size_t d) { return d > 1; }) != e) return false;
    for (size_t i = 0; i < e; ++i) {
        for (size_t k = rp[i]; k < rp[i + 1]; ++k) {
            const size_t j = ci[k];
            if (degree[j] > 1 && av[k] > 0.0) {
                if (state[i] != n || location[j] != e || av[k] != 4.0) return false;
                state[i] = j;
                location[j] = i;
            }
        }
        if (state[i] == n) return false;
    }
    for (size_t i = 0; i < e; ++i) {
        size_t count = 0;
        for (size_t k = rp[i]; k < rp[i + 1]; ++k) {
            const size_t j = location[ci[k]];
            if (j == e) continue;
            const bool nb = (i >= g && j == i - g) || (i + g < e && j == i + g) || (i % g && j == i - 1) ||
                            (i % g + 1 < g && j == i + 1);
            if (av[k] != (i == j ? 4.0 : (nb ? -1.0 : 0.0))) return false;
            ++count;
        }
        if (count != (size_t)(1 + (i >= g) + (i + g < e) + (i % g > 0) + (i % g + 1 < g))) return false;
    }
    std::vector<double> q(n, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t k = md->Q.row_ptr[i]; k < md->Q.row_ptr[i + 1]; ++k) {
            if (md->Q.col_ind[k] != i && md->Q.values[k] != 0.0) return false;
            if (md->Q.col_ind[k] == i) q[i] += md->Q.values[k];
        }
    for (size_t j : state)
        if (!(q[j] > 0.0) || !std::isfinite(q[j])) return false;
    return true;
}

}  // namespace hprqp
