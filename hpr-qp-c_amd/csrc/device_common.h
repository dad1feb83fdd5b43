// Shared device-side building blocks for the gfx950 HPR-QP kernels:
// error handling, deterministic wavefront/workgroup reductions, and the
// row-blocked CSR SpMV engine whose per-row results feed a fused epilogue.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <type_traits>

#include "shared/types.h"

namespace hprqp {

constexpr int kWave = 64;         // CDNA wavefront width
constexpr int kThreads = 256;     // workgroup size of every row/element kernel
constexpr int kStreamCap = 2048;  // nnz staged in LDS per stream-mode row block
constexpr int kLongRow = 256;     // rows longer than this leave the LDS stream path
constexpr int kMaxTerms = 8;      // reduction terms one kernel may emit

void report_hip_error(hipError_t err, const char *what, const char *file, int line);

// Status of a failed HIP call: a device allocation that ran out of memory is
// ALLOCATION_FAILED (as the reference reports a failed cudaMalloc,
// src/gpu/context.cu:89,163); every other device error NOT_IMPLEMENTED
// (src/gpu/cuda_common.h:52-55).
inline HPRQPStatus hip_status(hipError_t err)
{
    if (err == hipSuccess) return HPRQP_STATUS_OK;
    return err == hipErrorOutOfMemory ? HPRQP_STATUS_ALLOCATION_FAILED : HPRQP_STATUS_NOT_IMPLEMENTED;
}

#define HPRQP_HIP(call)                                                   \
    do {                                                                  \
        hipError_t hprqp_err_ = (call);                                   \
        if (hprqp_err_ != hipSuccess) {                                   \
            ::hprqp::report_hip_error(hprqp_err_, #call, __FILE__, __LINE__); \
            return ::hprqp::hip_status(hprqp_err_);                       \
        }                                                                 \
    } while (0)

// Device allocations and events owned by one scope: released on every return
// path of the host function that made them.
template <class T>
struct DeviceBuffer {
    T *p = nullptr;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) { return hipMalloc((void **)&p, sizeof(T) * (count ? count : 1)); }
};

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() = default;
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    hipError_t create()
    {
        hipError_t e = hipEventCreate(&a);
        return e != hipSuccess ? e : hipEventCreate(&b);
    }
};

// ---------------------------------------------------------------- reductions
// Fixed-shape trees, so results are run-to-run deterministic.

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

// The same sum through DPP lane moves, which stay in the vector ALU instead
// of the LDS crossbar that __shfl_down uses: pairs, quads, then rotations
// within each row of 16 lanes leave the row's sum in all of its lanes; the
// four rows are then read out and added, (r0 + r1) + (r2 + r3). The result
// is uniform over the wavefront. Every lane must be active.
template <int CTRL>
__device__ inline double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ inline double read_lane(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ inline double wave_sum_dpp(double v)
{
    v += dpp_move<0xb1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4e>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_move<0x124>(v);  // row_ror:4
    v += dpp_move<0x128>(v);  // row_ror:8
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

__device__ inline double wave_max(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, kWave));
    return v;
}

// Sum over the workgroup of NTH threads; the result is valid in every thread.
template <int NTH = kThreads>
__device__ inline double block_sum_all(double v)
{
    __shared__ double part[NTH / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) part[wid] = v;
    __syncthreads();
    double total = part[0];
#pragma unroll
    for (int w = 1; w < NTH / kWave; ++w) total += part[w];
    return total;
}

// Reduce NS sums and NM maxima over the workgroup of NTH threads and write
// them, in that order, to out[0 .. NS+NM) (threads 0 .. NS+NM-1 write).
template <int NS, int NM, int NTH = kThreads>
__device__ inline void block_reduce_store(const double *sums, const double *maxs, double *out)
{
    constexpr int NT = NS + NM;
    __shared__ double part[NTH / kWave][NT > 0 ? NT : 1];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        double v = wave_sum(sums[t]);
        if (lane == 0) part[wid][t] = v;
    }
#pragma unroll
    for (int t = 0; t < NM; ++t) {
        double v = wave_max(maxs[t]);
        if (lane == 0) part[wid][NS + t] = v;
    }
    __syncthreads();
    if (threadIdx.x < NT) {
        const int t = threadIdx.x;
        double acc = part[0][t];
        for (int w = 1; w < NTH / kWave; ++w) acc = t < NS ? acc + part[w][t] : fmax(acc, part[w][t]);
        out[t] = acc;
    }
}

// ------------------------------------------------------------------- SpMV
// Row blocks: block b covers rows [blk[b].x, blk[b+1].x) and entries
// [blk[b].y, blk[b+1].y), so one round trip gives a block its bounds. A block
// is one of three kinds, told apart by its mean row length:
//   * stream (mean <= kLongRow): up to kThreads rows of <= kLongRow entries,
//     <= kStreamCap entries in total. Products v*x[col] are staged through LDS
//     with coalesced loads; each row's owner thread then sums them left to
//     right without FMA, which reproduces the sequential row sum of the host
//     CSR product bit for bit.
//   * wave (rows longer than kLongRow, up to kThreads / kWave of them):
//     (kThreads / kWave) / rows wavefronts per row, each summing a 64-aligned
//     chunk with FMA over a 64-lane stride and a shuffle tree; the chunks'
//     sums are added in chunk order. A row longer than kVecRow gets a block
//     (all four wavefronts) of its own.
// Row bounds, carried partial sums and the epilogue's own row inputs are
// loaded before the gathers, so they overlap the longest latency chain.
//
// The epilogue type supplies:
//   static constexpr int NS, NM;              // reduction terms
//   __device__ void row(int r, double v, double *sums, double *maxs) const;
//   double *partials;                          // [gridDim.x][NS+NM] or null
// and optionally a prefetch split: `struct Pre`, `Pre load(int r) const` and
// `row_pre(int r, double v, const Pre &, double *sums, double *maxs) const`.
//
// Column panels: a matrix may be split into column panels (each its own CSR
// over the same rows). Panel p > 0 is run with CARRY_IN: a row's sum starts
// from the partial sum the previous panel left in `carry`, so with columns
// ascending inside each row the left-to-right order, and hence the bits, are
// those of the unsplit product; each panel gathers only a slice of x.

constexpr int kVecRow = 4096;  // rows longer than this get a block of their own

template <class E, class = void>
struct RowInputs {
    struct Pre {};
    __device__ static Pre load(const E &, int) { return Pre{}; }
    __device__ static void finish(const E &e, int r, double v, const Pre &, double *s, double *m) { e.row(r, v, s, m); }
};
template <class E>
struct RowInputs<E, std::void_t<typename E::Pre>> {
    using Pre = typename E::Pre;
    __device__ static Pre load(const E &e, int r) { return e.load(r); }
    __device__ static void finish(const E &e, int r, double v, const Pre &p, double *s, double *m)
    {
        e.row_pre(r, v, p, s, m);
    }
};

// Sum of one row's products over a wavefront (lane 0 holds the result).
__device__ inline double wave_row_sum(const int *__restrict__ ci, const double *__restrict__ val,
                                      const double *__restrict__ x, int k0, int k1, int lane)
{
    constexpr int U = 8;
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = 0.0;
    int k = k0 + lane;
    for (; k + (U - 1) * kWave < k1; k += U * kWave) {
        int c[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = __builtin_nontemporal_load(ci + k + u * kWave);
            v[u] = __builtin_nontemporal_load(val + k + u * kWave);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = fma(v[u], x[c[u]], a[u]);
    }
    // tail: fewer than U*kWave entries left; clamped, masked loads keep them in flight together
    if (k < k1) {
        int c[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k + u * kWave, k1 - 1);
            c[u] = __builtin_nontemporal_load(ci + kk);
            v[u] = __builtin_nontemporal_load(val + kk);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double p = v[u] * x[c[u]];
            if (k + u * kWave < k1) a[u] += p;
        }
    }
    const double t = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    return wave_sum(t);
}

template <class Epi, bool CARRY_IN = false>
__global__ __launch_bounds__(kThreads) void spmv_rows_kernel(const int *__restrict__ rp, const int *__restrict__ ci,
                                                             const double *__restrict__ val,
                                                             const double *__restrict__ x,
                                                             const int2 *__restrict__ blk, Epi epi,
                                                             const double *__restrict__ carry = nullptr)
{
    constexpr int NS = Epi::NS, NM = Epi::NM;
    using RI = RowInputs<Epi>;
    using Pre = typename RI::Pre;
    __shared__ double prod[kStreamCap];
    const int b = blockIdx.x, t = threadIdx.x;
    const int2 lo = blk[b], hi = blk[b + 1];
    const int r0 = lo.x, r1 = hi.x, k0 = lo.y, k1 = hi.y;
    const int nr = r1 - r0, cnt = k1 - k0;
    double sums[NS > 0 ? NS : 1];
    double maxs[NM > 0 ? NM : 1];
#pragma unroll
    for (int i = 0; i < (NS > 0 ? NS : 1); ++i) sums[i] = 0.0;
#pragma unroll
    for (int i = 0; i < (NM > 0 ? NM : 1); ++i) maxs[i] = 0.0;

    if (cnt > kLongRow * nr) {
        // wave mode: W = 4 / nr wavefronts per row (rows split into W
        // 64-aligned chunks whose partial sums are added in chunk order)
        __shared__ double part_sum[kThreads / kWave];
        const int lane = t & (kWave - 1), w = t / kWave;
        const int W = (kThreads / kWave) / nr;
        const int rl = w / W, part = w - rl * W;
        const int r = r0 + rl;
        Pre pre{};
        double cin = 0.0, acc = 0.0;
        if (rl < nr) {
            if (lane == 0 && part == 0) {
                pre = RI::load(epi, r);
                if (CARRY_IN) cin = carry[r];
            }
            const int rs = rp[r], re = rp[r + 1];
            const int chunk = (((re - rs) + W - 1) / W + kWave - 1) & ~(kWave - 1);
            const int ps = min(rs + part * chunk, re), pe = min(ps + chunk, re);
            acc = wave_row_sum(ci, val, x, ps, pe, lane);
        }
        if (W > 1) {
            if (lane == 0) part_sum[w] = acc;
            __syncthreads();
            if (lane == 0 && part == 0 && rl < nr) {
                acc = part_sum[w];
                for (int q = 1; q < W; ++q) acc += part_sum[w + q];
            }
        }
        if (rl < nr && lane == 0 && part == 0) {
            if (CARRY_IN) acc = cin + acc;
            RI::finish(epi, r, acc, pre, sums, maxs);
        }
    } else {
        const int r = r0 + t;
        Pre pre{};
        double acc = 0.0;
        int rs = 0, re = 0;
        if (r < r1) {
            pre = RI::load(epi, r);
            if (CARRY_IN) acc = carry[r];
            rs = rp[r] - k0;
            re = rp[r + 1] - k0;
        }
        constexpr int kPer = kStreamCap / kThreads;
        if (cnt > 0) {
            // Branch-free load phase: indices are clamped into the block so all
            // 2*kPer stream loads and kPer gathers issue before any use.
            int col[kPer];
            double v[kPer], xv[kPer];
            const int last = cnt - 1;
#pragma unroll
            for (int it = 0; it < kPer; ++it) {
                const int k = min(t + it * kThreads, last);
                col[it] = __builtin_nontemporal_load(ci + k0 + k);
                v[it] = __builtin_nontemporal_load(val + k0 + k);
            }
#pragma unroll
            for (int it = 0; it < kPer; ++it) xv[it] = x[col[it]];
#pragma unroll
            for (int it = 0; it < kPer; ++it) {
                const int k = t + it * kThreads;
                if (k < cnt) prod[k] = v[it] * xv[it];
            }
        }
        __syncthreads();
        if (r < r1) {
            for (int k = rs; k < re; ++k) acc += prod[k];
            RI::finish(epi, r, acc, pre, sums, maxs);
        }
    }
    if constexpr (NS + NM > 0) block_reduce_store<NS, NM>(sums, maxs, epi.partials + (size_t)b * (NS + NM));
}

// Elementwise map over [0, len) with optional reductions; one element per
// thread. The functor has the same interface as an SpMV epilogue minus the
// product value: row(int i, double *sums, double *maxs).
template <class F>
__global__ __launch_bounds__(kThreads) void map_kernel(int len, F f)
{
    constexpr int NS = F::NS, NM = F::NM;
    double sums[NS > 0 ? NS : 1];
    double maxs[NM > 0 ? NM : 1];
#pragma unroll
    for (int i = 0; i < (NS > 0 ? NS : 1); ++i) sums[i] = 0.0;
#pragma unroll
    for (int i = 0; i < (NM > 0 ? NM : 1); ++i) maxs[i] = 0.0;
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < len) f.row(i, sums, maxs);
    if constexpr (NS + NM > 0) block_reduce_store<NS, NM>(sums, maxs, f.partials + (size_t)blockIdx.x * (NS + NM));
}

// Finish a reduction: term t of out[] = reduce over nblk partial rows
// (stride `stride`), sum for t < ns, max otherwise. One workgroup per term.
__global__ __launch_bounds__(kThreads) void reduce_partials_kernel(const double *__restrict__ partials, int nblk,
                                                                   int stride, int ns, double *__restrict__ out);

}  // namespace hprqp
