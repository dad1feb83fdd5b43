// Batched HPR-QP: thousands of independent QPs solved by one persistent
// launch, one 256-thread workgroup per instance. Each workgroup owns its
// instance's whole state machine on the device -- Ruiz scaling, power
// iteration, the HPR iteration, residual checks, M-norm, restart and sigma
// updates -- so the host only relaunches the chunk kernel and reads a count of
// unfinished instances. The per-instance arithmetic is the plain solver's
// (same epilogue order, same control logic), so an instance's iterates match
// a plain solve of that instance up to the order of its reductions.
//
// No reference counterpart (the reference solves one QP per call); the loop it
// replaces is N sequential calls of solve_gpu (src/gpu/solver.cu:127-231).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>

#include <sys/mman.h>
#include <vector>

#include "device_common.h"
#include "hprqp.h"
#include "hprqp_hip.h"
#include "solve_loop.h"

namespace hprqp {
namespace batch {

// Per-instance layout in the concatenated device arrays.
struct Inst {
    int n, m;
    long long q_rp, q_nz;    // Q: row_ptr offset (n+1 entries), entry offset
    long long a_rp, a_nz;    // A: row_ptr offset (m+1), entry offset
    long long at_rp, at_nz;  // A^T: row_ptr offset (n+1), entry offset (perm into A entries)
    long long qt_rp;         // Q^T row_ptr offset (n+1); perm shares q_nz
    long long vn, vm;        // vector offsets
    int q_blk, q_nblk, a_blk, a_nblk, at_blk, at_nblk;  // row-block lists (offset, count)
    int dense;               // 1: Q, A held as dense row-major arrays in Arrays::dval
    long long dq, da, dat;   // dense offsets: Q (n x n), A^T (n x m), A (m x n)
    long long dg;            // dense: offset of G = A A^T (m x m) in Arrays::gram
    int noC;
    double norm_b_org, norm_c_org, obj_constant;
};

struct Restart {
    int flag, first;
    double last_gap, current_gap, save_gap;
    long long inner, times;
    double weighted_norm, best_gap, best_sigma;
};

// Device-resident solver state of one instance.
struct State {
    double sigma, lam_a, lam_q;
    int Q_is_diag, to_check, finished, status;
    long long iter, iter4, iter6;
    int first4, first6;
    double pobj, dobj, gap, err_rp, err_rd, kkt;
    Restart rs;
};

enum NV {
    N_x, N_x_bar, N_x_hat, N_w, N_w_bar, N_ATy, N_ATy_bar, N_ATdy, N_Qw, N_Qx, N_dQw, N_QATdy, N_dx, N_dw,
    N_z_bar, N_tempv, N_diag, N_fact1, N_fact2, N_fact, N_fact_M, N_last_x, N_last_w, N_last_ATy, N_c, N_l, N_u,
    N_col, N_tc, N_pq, N_pz, N_pt, N_Qw_bar, N_last_Qw, N_COUNT
};
enum MV { M_y, M_y_bar, M_dy, M_s, M_last_y, M_AL, M_AU, M_row, M_tr, M_pq, M_pz, M_e0, M_e1, M_e2, M_e3, M_COUNT };

struct Arrays {
    const Inst *inst;
    State *state;
    int *rp;          // all row pointers (Q, A, A^T, Q^T)
    int *ci;          // column indices: Q entries then A entries then A^T entries
    double *val;      // values: Q entries then A entries then A^T entries
    int *perm;        // A^T -> A entry (local), Q^T -> Q entry (local)
    int *blocks;      // row-block starts
    double *vn;       // N_COUNT vectors of total length sum_n, stacked
    double *vm;       // M_COUNT vectors of total length sum_m
    long long sum_n, sum_m;
    long long at_ci_base;  // where A^T column indices / values start in ci / val
    long long a_base;      // where A entries start in ci / val
    long long qt_perm_base;
    int *active;           // count of unfinished instances
    double *dval;          // dense instances' matrices (see Inst::dense)
    double *gram;          // dense instances' G = A A^T of the scaled A (setup only)
    const int *order;      // workgroup -> instance (largest first)
    int wave_row;          // CSR rows with at least this many entries get a wavefront
};

struct Params {
    double stoptol;
    long long max_iter;
    long long check_iter;
    int print_frequency;
    int use_scaling;
};

}  // namespace batch
}  // namespace hprqp

#include "epilogues.h"

namespace hprqp {
namespace batch {

// ------------------------------------------------------- workgroup engine

// y = M x over the instance's row blocks, each row's value handed to the
// epilogue (stream / long-row modes as in spmv_rows_kernel). Per-thread
// reduction terms accumulate across blocks; the caller reduces them.
// Rows of at least `wave_row` entries (chosen per batch by a timed probe)
// are grouped (up to 8 per block) and each is reduced by one wavefront (lanes
// stride the row, FMA, shuffle tree); the other rows go through the LDS
// stream mode with sequential row sums.
constexpr int kWaveRowDefault = 64;

// Matrix view of one instance. Sparse: CSR with row blocks. Dense (dv set):
// a row-major K x J array S whose product is taken in column form,
// y_j = sum_{i<K} S[i*ld + j] x_i, i ascending -- S = Q for Q (symmetric),
// S = A^T for A x, S = A for A^T y -- so every output is one thread's
// sequential sum, the order of the CSR row sum, with coalesced loads.
struct Mat {
    const int *rp, *ci;
    double *val;
    const int *blk;
    int nblk;
    double *dv;
    int K, J, ld;
    int wave_row;  // CSR blocks whose mean row length reaches this are wave blocks
};

// Largest dense dimension: x is staged whole in the LDS product buffer.
constexpr int kDenseMax = kStreamCap;
// CSR row blocks of the batched engine hold at most kBatchStreamCap entries
// (the products of a block are staged in LDS); the CSR solve kernel stages
// the gathered vector of each iteration product in LDS too when it has at
// most kStageMax entries, so that the gathers read LDS instead of missing L2
// again once the vector's lines are evicted.
constexpr int kBatchStreamCap = 1024;
constexpr int kStageMax = 2048;
constexpr int kLoneThreads = 1024;  // workgroup of an instance in the batch's tail

// Column-form dot of a dense view: sum_{i<K} S[i*ld + j] xs[i], i ascending,
// 32 loads in flight per thread. CACHED: ordinary loads, for products that
// re-read one instance's matrix many times (power iteration) -- a lone
// instance's matrices then stay in L2, where the loads in flight set its
// rate (16 in flight: the layer workload's power phase 69.3 ms, 32: 66.7 ms);
// otherwise streaming (non-temporal) loads, for the solve, whose matrices are
// read once per iteration (its dense kernel runs two waves per SIMD for the
// registers).
template <bool CACHED = false>
__device__ __forceinline__ double dense_column_dot(const Mat &M, int j, const double *xs)
{
    constexpr int U = 32;
    const double *__restrict__ col = M.dv + j;
    double acc = 0.0;
    int i = 0;
    for (; i + U <= M.K; i += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double *p = col + (long long)(i + u) * M.ld;
            v[u] = CACHED ? *p : __builtin_nontemporal_load(p);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u] * xs[i + u];
    }
    for (; i < M.K; ++i) acc += col[(long long)i * M.ld] * xs[i];
    return acc;
}

// Dense views run as "blocks" of kThreads outputs through the same epilogue
// call as the CSR stream mode, so the epilogue is inlined once per mode.
// NT > kThreads (CSR only): the workgroup runs as NT / kThreads groups of
// kThreads threads, group g taking row blocks g, g + NT / kThreads, ... with
// a staging area of its own (prod holds NT / kThreads areas); every row is
// summed exactly as with kThreads threads, only the threads that hold the
// epilogue's reduction terms change.
template <bool DENSE, bool CACHED = false, int NT = kThreads, class Epi>
__device__ __forceinline__ void wg_spmv(const Mat &M, const double *__restrict__ x, const Epi &epi, double *sums, double *maxs,
                        double *prod)
{
    constexpr bool dense = DENSE;
    static_assert(!DENSE || NT == kThreads, "dense views run with kThreads threads");
    const int *__restrict__ rp = M.rp;
    const int *__restrict__ ci = M.ci;
    const double *__restrict__ val = M.val;
    const int *__restrict__ blk = M.blk;
    if constexpr (NT > kThreads) {
        constexpr int NG = NT / kThreads;
        const int t = threadIdx.x % kThreads, grp = threadIdx.x / kThreads;
        const int lane = t & (kWave - 1), wid = t / kWave;
        double *gprod = prod + (size_t)grp * kBatchStreamCap;
        for (int b0 = 0; b0 < M.nblk; b0 += NG) {
            const int b = b0 + grp;
            const bool have = b < M.nblk;
            int r0 = 0, r1 = 0, k0 = 0, k1 = 0;
            if (have) {
                r0 = blk[b], r1 = blk[b + 1];
                k0 = rp[r0], k1 = rp[r1];
            }
            const bool wave_blk = have && (long long)(k1 - k0) >= (long long)M.wave_row * (r1 - r0);
            // every thread reaches both barriers, whatever its group's block kind
            if (wave_blk) {
                for (int r = r0 + wid; r < r1; r += kThreads / kWave) {
                    const int e = rp[r + 1];
                    double a0 = 0.0, a1 = 0.0;
                    int k = rp[r] + lane;
                    for (; k + kWave < e; k += 2 * kWave) {
                        a0 = fma(val[k], x[ci[k]], a0);
                        a1 = fma(val[k + kWave], x[ci[k + kWave]], a1);
                    }
                    if (k < e) a0 = fma(val[k], x[ci[k]], a0);
                    const double acc = wave_sum(a0 + a1);
                    if (lane == 0) epi.row(r, acc, sums, maxs);
                }
            } else if (have) {
                const int cnt = k1 - k0;
                for (int k = t; k < cnt; k += kThreads) gprod[k] = val[k0 + k] * x[ci[k0 + k]];
            }
            __syncthreads();
            const int r = r0 + t;
            if (have && !wave_blk && r < r1) {
                double acc = 0.0;
                const int e = rp[r + 1] - k0;
                for (int k = rp[r] - k0; k < e; ++k) acc += gprod[k];
                epi.row(r, acc, sums, maxs);
            }
            __syncthreads();
        }
        __syncthreads();
        return;
    }
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1), wid = t / kWave;
    int nblk = M.nblk;
    if constexpr (dense) {
        // x staged whole in LDS (K <= kDenseMax)
        for (int i = t; i < M.K; i += kThreads) prod[i] = x[i];
        __syncthreads();
        nblk = (M.J + kThreads - 1) / kThreads;
    }
    for (int b = 0; b < nblk; ++b) {
        int r0, r1, k0 = 0, k1 = 0;
        if constexpr (dense) {
            r0 = b * kThreads;
            r1 = min(r0 + kThreads, M.J);
        } else {
            r0 = blk[b], r1 = blk[b + 1];
            k0 = rp[r0], k1 = rp[r1];
        }
        if (!dense && (long long)(k1 - k0) >= (long long)M.wave_row * (r1 - r0)) {
            for (int r = r0 + wid; r < r1; r += kThreads / kWave) {
                const int e = rp[r + 1];
                double a0 = 0.0, a1 = 0.0;
                int k = rp[r] + lane;
                for (; k + kWave < e; k += 2 * kWave) {
                    a0 = fma(val[k], x[ci[k]], a0);
                    a1 = fma(val[k + kWave], x[ci[k + kWave]], a1);
                }
                if (k < e) a0 = fma(val[k], x[ci[k]], a0);
                const double acc = wave_sum(a0 + a1);
                if (lane == 0) epi.row(r, acc, sums, maxs);
            }
        } else {
            const int cnt = k1 - k0;
            if (!dense) {
                for (int k = t; k < cnt; k += kThreads) prod[k] = val[k0 + k] * x[ci[k0 + k]];
                __syncthreads();
            }
            const int r = r0 + t;
            if (r < r1) {
                double acc = 0.0;
                if constexpr (dense) {
                    acc = dense_column_dot<CACHED>(M, r, prod);
                } else {
                    const int e = rp[r + 1] - k0;
                    for (int k = rp[r] - k0; k < e; ++k) acc += prod[k];
                }
                epi.row(r, acc, sums, maxs);
            }
        }
        if (!dense) __syncthreads();
    }
    __syncthreads();
}

// An iteration product. A staging workgroup (CSR, every vector of the batch
// at most kStageMax long) copies the gathered vector x to LDS first and the
// product gathers from there: the same sums in the same order.
template <class WG, class Epi>
__device__ __forceinline__ void wg_product(WG &g, const Mat &M, const double *x, int xlen, const Epi &epi, double *sums,
                           double *maxs)
{
    if constexpr (WG::kStage) {
        for (int i = threadIdx.x; i < xlen; i += WG::kNT) g.xs[i] = x[i];
        __syncthreads();
        wg_spmv<false, false, WG::kNT>(M, g.xs, epi, sums, maxs, g.prod);
    } else {
        wg_spmv<WG::kDense, false, WG::kNT>(M, x, epi, sums, maxs, g.prod);
    }
}

template <int NT = kThreads, class F>
__device__ __forceinline__ void wg_map(int len, const F &f, double *sums, double *maxs)
{
    for (int i = threadIdx.x; i < len; i += NT) f.row(i, sums, maxs);
    __syncthreads();
}

// Reduce per-thread terms; every thread then reads out[0 .. NS+NM).
template <int NS, int NM, int NT = kThreads>
__device__ __forceinline__ void wg_reduce(double *sums, double *maxs, double *out)
{
    block_reduce_store<NS, NM, NT>(sums, maxs, out);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < (NS > 0 ? NS : 1); ++i) sums[i] = 0.0;
#pragma unroll
    for (int i = 0; i < (NM > 0 ? NM : 1); ++i) maxs[i] = 0.0;
}

// DiagInfo on a dense row-major Q.
struct DenseDiagInfo {
    static constexpr int NS = 0, NM = 1;
    double *partials;
    double *diag;
    const double *Qd;
    int n;
    __device__ void row(int r, double *, double *mx) const
    {
        double off = 0.0;
        for (int c = 0; c < n; ++c)
            if (c != r) off = fmax(off, fabs(Qd[(long long)r * n + c]));
        diag[r] = 0.0 + Qd[(long long)r * n + r];
        mx[0] = fmax(mx[0], off);
    }
};

__device__ __forceinline__ double dclamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }


// One workgroup's view of its instance; D: the instance's matrices are dense;
// ST: the iteration products stage their gathered vector in LDS (xs).
template <bool D, bool ST = false, int NT = kThreads>
struct WGv {
    static constexpr bool kDense = D;
    static constexpr bool kStage = ST;
    static constexpr int kNT = NT;  // threads of the workgroup
    const Inst *I;
    State *S;  // in LDS
    Mat Q, A, AT;
    Mat G;  // dense: A A^T (m x m) of the scaled A, for the lambda_max(A A^T) estimate
    const int *at_perm, *qt_rp, *qt_perm;
    double *a_val;
    double *vn[N_COUNT];
    double *vm[M_COUNT];
    double *prod;   // LDS staging
    double *red;    // LDS reduction results
    double *xs = nullptr;  // ST: LDS copy of a product's gathered vector (<= kStageMax entries)
    Params P;

    __device__ double *N(int v) const { return vn[v]; }
    __device__ double *M(int v) const { return vm[v]; }
};

// ---------------------------------------------------------- setup phases

// The same passes on a dense instance (Q: n x n, A: m x n, both row-major;
// the A^T copy is made after scaling). Zero entries leave every max and sum
// unchanged, so the factors equal the sparse path's.
template <class WG>
__device__ void wg_scale_dense(WG &g)
{
    const int n = g.I->n, m = g.I->m, t = threadIdx.x;
    double *Qd = g.Q.dv, *Ad = g.AT.dv;  // AT's column-form array is A row-major
    for (int pass = 0; pass < 11; ++pass) {
        const bool sum = pass == 10;
        for (int j = t; j < n; j += kThreads) {
            double acc = 0.0;
            for (int i = 0; i < m; ++i) {
                const double v = fabs(Ad[(long long)i * n + j]);
                if (sum) acc += v;
                else if (v > acc) acc = v;
            }
            for (int i = 0; i < n; ++i) {
                const double v = fabs(Qd[(long long)i * n + j]);
                if (sum) acc += v;
                else if (v > acc) acc = v;
            }
            acc = sqrt(acc);
            g.N(N_tc)[j] = acc == 0.0 ? 1.0 : acc;
        }
        if (sum) {
            // the sum in entry order, one thread per row
            for (int i = t; i < m; i += kThreads) {
                double acc = 0.0;
                for (int j = 0; j < n; ++j) acc += fabs(Ad[(long long)i * n + j]);
                acc = sqrt(acc);
                g.M(M_tr)[i] = acc == 0.0 ? 1.0 : acc;
            }
        } else {
            // the max (order-free) by one wavefront per row: coalesced loads
            const int lane = t & (kWave - 1);
            for (int i = t / kWave; i < m; i += kThreads / kWave) {
                double acc = 0.0;
                for (int j = lane; j < n; j += kWave) acc = fmax(acc, fabs(Ad[(long long)i * n + j]));
                acc = sqrt(wave_max(acc));
                if (lane == 0) g.M(M_tr)[i] = acc == 0.0 ? 1.0 : acc;
            }
        }
        __syncthreads();
        const double *tc = g.N(N_tc), *tr = g.M(M_tr);
        for (long long e = t; e < (long long)n * n; e += kThreads) {
            const int r = (int)(e / n), c = (int)(e % n);
            double v = Qd[e];
            v /= tc[r];
            v /= tc[c];
            Qd[e] = v;
        }
        for (long long e = t; e < (long long)m * n; e += kThreads) {
            const int r = (int)(e / n), c = (int)(e % n);
            double v = Ad[e];
            v /= tr[r];
            v /= tc[c];
            Ad[e] = v;
        }
        for (int r = t; r < n; r += kThreads) {
            const double d = tc[r];
            g.N(N_col)[r] *= d;
            g.N(N_c)[r] /= d;
            g.N(N_l)[r] *= d;
            g.N(N_u)[r] *= d;
        }
        for (int r = t; r < m; r += kThreads) {
            const double d = tr[r];
            g.M(M_row)[r] *= d;
            if (!isinf(g.M(M_AL)[r])) g.M(M_AL)[r] /= d;
            if (!isinf(g.M(M_AU)[r])) g.M(M_AU)[r] /= d;
        }
        __syncthreads();
    }
}

// scale.c:231-318 on one instance (column j of A through A^T's permutation,
// column j of Q through Q^T's, so sums run in row order like the host).
template <class WG>
__device__ void wg_scale(WG &g)
{
    const int n = g.I->n, m = g.I->m, t = threadIdx.x;
    for (int j = t; j < n; j += kThreads) g.N(N_col)[j] = 1.0;
    for (int i = t; i < m; i += kThreads) g.M(M_row)[i] = 1.0;
    __syncthreads();
    if (!g.P.use_scaling) return;
    if constexpr (WG::kDense) {
        wg_scale_dense(g);
        return;
    }
    for (int pass = 0; pass < 11; ++pass) {
        const bool sum = pass == 10;
        for (int j = t; j < n; j += kThreads) {
            double acc = 0.0;
            for (int k = g.AT.rp[j]; k < g.AT.rp[j + 1]; ++k) {
                const double v = fabs(g.a_val[g.at_perm[k]]);
                if (sum) acc += v;
                else if (v > acc) acc = v;
            }
            for (int k = g.qt_rp[j]; k < g.qt_rp[j + 1]; ++k) {
                const double v = fabs(g.Q.val[g.qt_perm[k]]);
                if (sum) acc += v;
                else if (v > acc) acc = v;
            }
            acc = sqrt(acc);
            g.N(N_tc)[j] = acc == 0.0 ? 1.0 : acc;
        }
        for (int i = t; i < m; i += kThreads) {
            double acc = 0.0;
            for (int k = g.A.rp[i]; k < g.A.rp[i + 1]; ++k) {
                const double v = fabs(g.A.val[k]);
                if (sum) acc += v;
                else if (v > acc) acc = v;
            }
            acc = sqrt(acc);
            g.M(M_tr)[i] = acc == 0.0 ? 1.0 : acc;
        }
        __syncthreads();
        const double *tc = g.N(N_tc), *tr = g.M(M_tr);
        for (int r = t; r < n; r += kThreads) {
            const double d = tc[r];
            for (int k = g.Q.rp[r]; k < g.Q.rp[r + 1]; ++k) {
                double v = g.Q.val[k];
                v /= d;
                v /= tc[g.Q.ci[k]];
                g.Q.val[k] = v;
            }
            g.N(N_col)[r] *= d;
            g.N(N_c)[r] /= d;
            g.N(N_l)[r] *= d;
            g.N(N_u)[r] *= d;
        }
        for (int r = t; r < m; r += kThreads) {
            const double d = tr[r];
            for (int k = g.A.rp[r]; k < g.A.rp[r + 1]; ++k) {
                double v = g.A.val[k];
                v /= d;
                v /= tc[g.A.ci[k]];
                g.A.val[k] = v;
            }
            g.M(M_row)[r] *= d;
            if (!isinf(g.M(M_AL)[r])) g.M(M_AL)[r] /= d;
            if (!isinf(g.M(M_AU)[r])) g.M(M_AU)[r] /= d;
        }
        __syncthreads();
    }
}

// eigs.c:19-153 on one instance (a_matrix: A A^T on m-vectors; else Q).
template <class WG>
__device__ double wg_power(WG &g, bool a_matrix, double *sums, double *maxs)
{
    const int len = a_matrix ? g.I->m : g.I->n;
    if (len == 0) return 0.0;
    double *z = a_matrix ? g.M(M_pz) : g.N(N_pz);
    double *q = a_matrix ? g.M(M_pq) : g.N(N_pq);
    double *tmp = g.N(N_pt);
    wg_map(len, MapProbe{nullptr, z, 0.0, (double)len + 1.0}, sums, maxs);
    wg_reduce<1, 0>(sums, maxs, g.red);
    double qn = sqrt(g.red[0]);
    double lam = 1.0;
    for (int it = 0; it < 5000; ++it) {
        if (qn <= 1e-16) {
            lam = 0.0;
            break;
        }
        // later steps' q = z / qn comes out of the previous residual pass
        if (it == 0) wg_map(len, MapScaleDiv{{}, q, z, qn}, sums, maxs);
        if (a_matrix) {
            if constexpr (WG::kDense) {
                // z = G q with G = A A^T formed once: the same iteration as
                // A (A^T q) up to rounding, one matrix pass instead of two
                (void)tmp;
                wg_spmv<true, true>(g.G, q, EpiPowerZ{nullptr, z, q}, sums, maxs, g.prod);
            } else {
                wg_spmv<false, true>(g.AT, q, EpiStore{{}, tmp}, sums, maxs, g.prod);
                wg_spmv<false, true>(g.A, tmp, EpiPowerZ{nullptr, z, q}, sums, maxs, g.prod);
            }
        } else {
            wg_spmv<WG::kDense, true>(g.Q, q, EpiPowerZ{nullptr, z, q}, sums, maxs, g.prod);
        }
        wg_reduce<2, 0>(sums, maxs, g.red);
        lam = g.red[0];
        const double zn = sqrt(g.red[1]);
        __syncthreads();
        wg_map(len, MapPowerResid{nullptr, q, z, g.red}, sums, maxs);
        wg_reduce<1, 0>(sums, maxs, g.red + 2);
        const double err = sqrt(g.red[2]) / (zn + fabs(lam) + 1e-16);
        __syncthreads();
        if (err < 1e-4) break;
        qn = zn;
    }
    return lam < 0.0 ? 0.0 : lam;
}

// workspace.c:191-248, in two parts: scaling, the A^T (or dense
// transposed) copy, sigma0 and the diagonal-Q test; then the power iterations.
// A dense batch forms G = A A^T between the two (dense_gram_kernel).
template <class WG>
__device__ void wg_setup_pre(WG &g, double *sums, double *maxs)
{
    const int n = g.I->n, m = g.I->m, t = threadIdx.x;
    wg_scale(g);
    if constexpr (WG::kDense) {
        // A x runs in column form over A^T: copy the scaled A transposed
        for (long long e = t; e < (long long)m * n; e += kThreads) {
            const int i = (int)(e / n), j = (int)(e % n);
            g.A.dv[(long long)j * m + i] = g.AT.dv[e];
        }
    } else {
        for (long long k = t; k < (long long)g.AT.rp[n]; k += kThreads) g.AT.val[k] = g.a_val[g.at_perm[k]];
    }
    __syncthreads();
    // sigma0 = ||b||/||c|| on the scaled data
    wg_map(n, MapDot{nullptr, g.N(N_c), g.N(N_c)}, sums, maxs);
    wg_reduce<1, 0>(sums, maxs, g.red);
    const double nc = sqrt(g.red[0]);
    for (int i = t; i < m; i += kThreads) {
        const double al = isinf(g.M(M_AL)[i]) ? 0.0 : fabs(g.M(M_AL)[i]);
        const double au = isinf(g.M(M_AU)[i]) ? 0.0 : fabs(g.M(M_AU)[i]);
        const double v = al > au ? al : au;
        sums[0] += v * v;
    }
    __syncthreads();
    wg_reduce<1, 0>(sums, maxs, g.red + 1);
    const double nb = sqrt(g.red[1]);
    __syncthreads();
    double sigma = (nb > 1e-16 && nc > 1e-16 && nb < 1e16 && nc < 1e16) ? nb / nc : 1.0;
    if constexpr (WG::kDense)
        wg_map(n, DenseDiagInfo{nullptr, g.N(N_diag), g.Q.dv, n}, sums, maxs);
    else
        wg_map(n, DiagInfo{nullptr, g.N(N_diag), g.Q.rp, g.Q.ci, g.Q.val}, sums, maxs);
    wg_reduce<0, 1>(sums, maxs, g.red);
    const int diag = n == 0 || g.red[0] <= 1e-15;
    __syncthreads();
    if (diag)
        wg_map(n, MapDiagFactors{{}, g.N(N_fact1), g.N(N_fact2), g.N(N_fact), g.N(N_fact_M), g.N(N_diag), sigma},
               sums, maxs);
    if (t == 0) {
        g.S->sigma = sigma;
        g.S->Q_is_diag = diag;
    }
    __syncthreads();
}

// lambdas_given: batch_power_kernel has stored lambda_max(A A') and, for a
// non-diagonal Q, lambda_max(Q) (both x 1.05) in the state.
template <class WG>
__device__ void wg_setup_post(WG &g, double *sums, double *maxs, bool lambdas_given = false)
{
    const int n = g.I->n, t = threadIdx.x;
    const int diag = g.S->Q_is_diag;
    double lam_a = lambdas_given ? g.S->lam_a : wg_power(g, true, sums, maxs) * 1.05;
    double lam_q = 0.0;
    if (lambdas_given && !diag) {
        lam_q = g.S->lam_q;
    } else if (diag) {
        for (int i = t; i < n; i += kThreads) maxs[0] = fmax(maxs[0], g.N(N_diag)[i]);
        wg_reduce<0, 1>(sums, maxs, g.red);
        lam_q = g.red[0];
        __syncthreads();
    } else {
        lam_q = wg_power(g, false, sums, maxs) * 1.05;
    }
    if (t == 0) {
        g.S->lam_a = lam_a;
        g.S->lam_q = lam_q;
        g.S->Q_is_diag = diag;
        g.S->to_check = 1;
    }
    __syncthreads();
}

}  // namespace batch
}  // namespace hprqp

namespace hprqp {
namespace batch {

// ------------------------------------------------------ control (thread 0)
// src/cpu/control.c:231-305, 45-77, 385-404 and solver_backend.c:119-171.

__device__ void d_restart_check(Restart &r, long long iter, long long check_iter, double sigma)
{
    if (check_iter == 0) return;
    r.flag = 0;
    if (r.first) {
        if (iter == check_iter) {
            r.first = 0;
            r.flag = 1;
            r.weighted_norm = r.current_gap;
            r.best_gap = r.current_gap;
            r.best_sigma = sigma;
        }
        return;
    }
    if (iter % check_iter != 0) return;
    if (r.current_gap < 0.0) r.current_gap = 1e-6;
    if (r.current_gap <= 0.2 * r.last_gap) r.flag = 1;
    if (r.current_gap <= 0.8 * r.last_gap && r.current_gap > r.save_gap) r.flag = 2;
    const double fact = r.weighted_norm > 0.0 && r.current_gap / r.weighted_norm > 1e-1 ? 0.5 : 0.2;
    if ((double)r.inner >= fact * (double)iter) r.flag = 3;
    if (r.best_gap > r.current_gap) {
        r.best_gap = r.current_gap;
        r.best_sigma = sigma;
    }
    r.save_gap = r.current_gap;
}

__device__ void d_restart_post(Restart &r, long long iter, long long check_iter, double gap)
{
    const bool end = check_iter > 0 && (iter + 1) % check_iter == 0;
    if (r.flag > 0 || end) {
        if (r.flag > 0) r.last_gap = gap;
        if (end) r.current_gap = gap;
    }
}

__device__ double d_sigma_obj(double x, double a, double b, double c, double d)
{
    return a * x + b / x + c * x * x / (1.0 + d * x);
}

__device__ double d_golden(double a, double b, double c, double d)
{
    double lo = 1e-12, hi = 1e12;
    const double g = (sqrt(5.0) - 1.0) / 2.0, g2 = (3.0 - sqrt(5.0)) / 2.0;
    double x1 = lo + g2 * (hi - lo), x2 = lo + g * (hi - lo);
    double f1 = d_sigma_obj(x1, a, b, c, d), f2 = d_sigma_obj(x2, a, b, c, d);
    for (int it = 0; it < 256; ++it) {
        if (fabs(hi - lo) <= 1e-13 * fmax(1.0, fmax(fabs(lo), fabs(hi)))) break;
        if (f1 > f2) {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + g * (hi - lo);
            f2 = d_sigma_obj(x2, a, b, c, d);
        } else {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = lo + g2 * (hi - lo);
            f1 = d_sigma_obj(x1, a, b, c, d);
        }
    }
    return 0.5 * (lo + hi);
}

__device__ double d_sigma_from_estimate(double est, double old, const Restart &r, const State &s)
{
    est = dclamp(est, 1e-12, 1e12);
    const double fact = (r.best_gap > 0.0 && isfinite(r.best_gap)) ? exp(-0.05 * (r.current_gap / r.best_gap)) : 1.0;
    const double cand = exp(fact * log(est) + (1.0 - fact) * log(r.best_sigma > 0.0 ? r.best_sigma : old));
    const double t1 = fmax(fmin(s.err_rd, s.err_rp), fmin(s.gap, r.current_gap));
    double kappa = 1.0;
    if (t1 <= 9e-10 && s.err_rp > 0.0) {
        const double ratio = s.err_rd / s.err_rp;
        kappa = t1 > 5e-10 ? dclamp(sqrt(ratio), 1e-2, 100.0) : dclamp(ratio, 1e-2, 100.0);
    }
    return dclamp(kappa * cand, 1e-12, 1e12);
}

__device__ long long d_print_step(long long iter)
{
    long long step = 10;
    if (iter == 0) return step;
    while (iter / step >= 100) step *= 10;
    return step;
}

__device__ bool d_print_now(const Params &p, long long iter)
{
    if (p.print_frequency == -1 && (iter % d_print_step(iter) == 0 || iter == p.max_iter)) return true;
    if (p.print_frequency > 0 && (iter % p.print_frequency == 0 || iter == p.max_iter)) return true;
    return false;
}

__device__ int d_check_next(const Params &p, const Restart &r, long long next)
{
    int chk = 0;
    if (p.check_iter > 0 && next % p.check_iter == 0) chk = 1;
    if (r.flag > 0) chk = 1;
    if (p.print_frequency == -1 && next % d_print_step(next) == 0) chk = 1;
    if (p.print_frequency > 0 && next % p.print_frequency == 0) chk = 1;
    return chk;
}

// ----------------------------------------------------------- operators

// residuals.c:88-189 (CPU convention for the dual objective)
template <class WG>
__device__ __forceinline__ void wg_residuals(WG &g, long long iter, double *sums, double *maxs)
{
    const int n = g.I->n, m = g.I->m;
    State &S = *g.S;
    wg_spmv<WG::kDense, false, WG::kNT>(g.Q, g.N(N_x_bar), EpiResQ{nullptr, g.N(N_Qx), g.N(N_x_bar), g.N(N_c), g.N(N_z_bar)}, sums, maxs, g.prod);
    wg_reduce<3, 0, WG::kNT>(sums, maxs, g.red);
    const double cx = g.red[0], xqx = g.red[1], zx = g.red[2];
    __syncthreads();
    wg_spmv<WG::kDense, false, WG::kNT>(g.AT, g.M(M_y_bar), EpiResAT{nullptr, g.N(N_ATdy), g.N(N_Qx), g.N(N_c), g.N(N_z_bar), g.N(N_col), 1.0}, sums, maxs, g.prod);
    wg_reduce<0, 3, WG::kNT>(sums, maxs, g.red);
    const double rd = g.red[0], qxm = g.red[1], atym = g.red[2];
    __syncthreads();
    double ys = 0.0, rp = 0.0, axm = 0.0;
    if (m > 0) {
        wg_spmv<WG::kDense, false, WG::kNT>(g.A, g.N(N_x_bar), EpiResA{nullptr, g.M(M_pz), g.M(M_AL), g.M(M_AU), g.M(M_row), g.M(M_y_bar), g.M(M_s), 1.0}, sums,
                maxs, g.prod);
        wg_reduce<1, 2, WG::kNT>(sums, maxs, g.red);
        ys = g.red[0];
        rp = g.red[1];
        axm = g.red[2];
        __syncthreads();
    }
    double viol = 0.0;
    if (iter == 0) {
        wg_map<WG::kNT>(n, MapBoundViolation{nullptr, g.N(N_dx), g.N(N_l), g.N(N_u), g.N(N_x_bar), g.N(N_col), 1.0}, sums,
               maxs);
        wg_reduce<0, 1, WG::kNT>(sums, maxs, g.red);
        viol = g.red[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        S.pobj = (cx + 0.5 * xqx) + g.I->obj_constant;
        S.dobj = (-0.5 * xqx + zx + ys) + g.I->obj_constant;
        S.gap = fabs(S.pobj - S.dobj) / (1.0 + fmax(fabs(S.pobj), fabs(S.dobj)));
        S.err_rd = rd / (1.0 + fmax(fmax(g.I->norm_c_org, atym), qxm));
        S.err_rp = m > 0 ? rp / (1.0 + fmax(g.I->norm_b_org, axm)) : 0.0;
        if (iter == 0 && viol > S.err_rp) S.err_rp = viol;
        S.kkt = fmax(fmax(S.err_rp, S.err_rd), S.gap);
    }
    __syncthreads();
}

// control.c:155-229
template <class WG>
__device__ __forceinline__ double wg_m_norm(WG &g, double *sums, double *maxs)
{
    const int m = g.I->m;
    const State &S = *g.S;
    const int diag = S.Q_is_diag;
    wg_spmv<WG::kDense, false, WG::kNT>(g.Q, g.N(N_dw), EpiMnQdw{nullptr, g.N(N_dQw), g.N(N_dx), g.N(N_dw)},
            sums, maxs, g.prod);
    wg_reduce<4, 0, WG::kNT>(sums, maxs, g.red);
    const double dx_sq = g.red[0], dqw_dx = g.red[1], dw_dqw = g.red[2], dqw_sq = g.red[3];
    __syncthreads();
    const double sg = S.sigma;
    double m1 = 0.0, m3 = 0.0, m2 = (1.0 / sg) * dx_sq;
    m2 -= 2.0 * dqw_dx;
    if (m > 0) {
        wg_spmv<WG::kDense, false, WG::kNT>(g.AT, g.M(M_dy), EpiMnATdy{nullptr, g.N(N_ATdy), g.N(N_dx)},
                sums, maxs, g.prod);
        wg_reduce<1, 0, WG::kNT>(sums, maxs, g.red);
        const double atdy_dx = g.red[0];
        __syncthreads();
        wg_spmv<WG::kDense, false, WG::kNT>(g.Q, g.N(N_ATdy), EpiMnQATdy{nullptr, g.N(N_QATdy), g.N(N_dQw), g.N(N_ATdy), g.N(N_fact_M), diag}, sums, maxs, g.prod);
        wg_reduce<3, 0, WG::kNT>(sums, maxs, g.red);
        const double dqw_atdy = g.red[0], atdy_qatdy = g.red[1], fm_term = g.red[2];
        __syncthreads();
        wg_map<WG::kNT>(m, MapDot{nullptr, g.M(M_dy), g.M(M_dy)}, sums, maxs);
        wg_reduce<1, 0, WG::kNT>(sums, maxs, g.red);
        const double dy_sq = g.red[0];
        __syncthreads();
        m1 = sg * S.lam_a * dy_sq;
        m1 -= 2.0 * sg * dqw_atdy;
        m2 += 2.0 * atdy_dx;
        if (diag) {
            m3 = fm_term;
            m1 += sg * dqw_sq;
        } else {
            m3 = (sg * sg) / (1.0 + sg * S.lam_q) * atdy_qatdy;
            m1 += sg * S.lam_q * dw_dqw;
        }
    } else if (!diag) {
        m1 = sg * S.lam_q * dw_dqw;
    }
    m2 += m1 > 0.0 ? m1 : 0.0;
    return sqrt((m2 > 0.0 ? m2 : 0.0) + (m3 > 0.0 ? m3 : 0.0));
}

// control.c:307-411 (diagonal Q: golden search over the workgroup)
template <class WG>
__device__ __forceinline__ double wg_golden_diag(WG &g, double a, double b)
{
    const int n = g.I->n;
    const double *dq = g.N(N_diag), *c = g.N(N_ATdy), *d = g.N(N_QATdy);
    auto eval = [&](double x) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < n; i += WG::kNT) acc += c[i] * (d[i] / (1.0 + x * dq[i]));
        const double tot = block_sum_all<WG::kNT>(acc);
        return a * x + b / x + x * x * tot;
    };
    double lo = 1e-12, hi = 1e12;
    const double phi = (sqrt(5.0) - 1.0) / 2.0;
    double x1 = hi - phi * (hi - lo), x2 = lo + phi * (hi - lo);
    double f1 = eval(x1), f2 = eval(x2);
    int it = 0;
    while (fabs(hi - lo) > 1e-12 * fmax(1.0, fabs(lo)) && it < 200) {
        if (f1 > f2) {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + phi * (hi - lo);
            f2 = eval(x2);
        } else {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = hi - phi * (hi - lo);
            f1 = eval(x1);
        }
        ++it;
    }
    return f1 < f2 ? x1 : x2;
}

template <class WG>
__device__ __forceinline__ void wg_update_sigma(WG &g, double *sums, double *maxs)
{
    State &S = *g.S;
    if (S.rs.flag < 1) return;
    const int n = g.I->n, m = g.I->m;
    const int diag = S.Q_is_diag;
    const double old = S.sigma;
    wg_map<WG::kNT>(n, MapSigmaDiffN{nullptr, g.N(N_dx), g.N(N_dw), g.N(N_ATdy), g.N(N_x_bar), g.N(N_last_x), g.N(N_w_bar),
                            g.N(N_last_w), g.N(N_ATy_bar), g.N(N_last_ATy), m > 0 ? 1 : 0},
           sums, maxs);
    wg_reduce<1, 0, WG::kNT>(sums, maxs, g.red);
    double b = g.red[0];
    __syncthreads();
    double dy_sq = 0.0;
    if (m > 0) {
        wg_map<WG::kNT>(m, MapSigmaDiffM{nullptr, g.M(M_dy), g.M(M_y_bar), g.M(M_last_y)}, sums, maxs);
        wg_reduce<1, 0, WG::kNT>(sums, maxs, g.red);
        dy_sq = g.red[0];
        __syncthreads();
    }
    wg_spmv<WG::kDense, false, WG::kNT>(g.Q, g.N(N_dw), EpiSigmaQdw{nullptr, g.N(N_dQw), g.N(N_ATdy), g.N(N_dw)}, sums, maxs, g.prod);
    wg_reduce<3, 0, WG::kNT>(sums, maxs, g.red);
    const double dqw_atdy = g.red[0], dqw_sq = g.red[1], dw_dqw = g.red[2];
    __syncthreads();
    double atdy_qatdy = 0.0;
    if (m > 0) {
        wg_spmv<WG::kDense, false, WG::kNT>(g.Q, g.N(N_ATdy), EpiSigmaQATdy{nullptr, g.N(N_QATdy), g.N(N_ATdy)}, sums, maxs, g.prod);
        wg_reduce<1, 0, WG::kNT>(sums, maxs, g.red);
        atdy_qatdy = g.red[0];
        __syncthreads();
    }
    double a = 0.0, cterm = 0.0, dterm = 0.0;
    if (m > 0) {
        a = S.lam_a * dy_sq;
        a -= 2.0 * dqw_atdy;
    }
    if (diag) {
        a += dqw_sq;
    } else {
        a += S.lam_q * dw_dqw;
        if (m > 0) {
            cterm = atdy_qatdy;
            dterm = S.lam_q;
        }
    }
    a = a > 1e-12 ? a : 1e-12;
    b = b > 1e-12 ? b : 1e-12;
    double est;
    if (diag && m > 0) est = wg_golden_diag(g, a, b);
    else if (!diag && m > 0) est = d_golden(a, b, cterm > 0.0 ? cterm : 0.0, dterm > 0.0 ? dterm : 0.0);
    else est = sqrt(b / a);
    const double sg = d_sigma_from_estimate(est, old, S.rs, S);
    __syncthreads();
    if (threadIdx.x == 0) S.sigma = sg;
    if (diag && fabs(sg - old) > 1e-15)
        wg_map<WG::kNT>(n, MapDiagFactors{{}, g.N(N_fact1), g.N(N_fact2), g.N(N_fact), g.N(N_fact_M), g.N(N_diag), sg}, sums,
               maxs);
    __syncthreads();
}

// control.c:413-457
template <class WG>
__device__ __forceinline__ void wg_restart_apply(WG &g, double *sums, double *maxs)
{
    State &S = *g.S;
    if (S.rs.flag <= 0) return;
    const int m = g.I->m;
    wg_spmv<WG::kDense, false, WG::kNT>(g.AT, g.M(M_y_bar), EpiRestart{{}, g.N(N_ATy_bar), g.N(N_ATy), g.N(N_last_ATy), g.N(N_x), g.N(N_w), g.N(N_last_x),
                       g.N(N_last_w), g.N(N_Qw), g.N(N_last_Qw), g.N(N_x_bar), g.N(N_w_bar), g.N(N_Qw_bar),
                       m > 0 ? 1 : 0, g.I->noC},
            sums, maxs, g.prod);
    wg_map<WG::kNT>(m, MapRestartM{{}, g.M(M_y), g.M(M_last_y), g.M(M_y_bar)}, sums, maxs);
    if (threadIdx.x == 0) {
        S.rs.last_gap = S.rs.current_gap;
        S.rs.save_gap = INFINITY;
        S.rs.times += 1;
        S.rs.inner = 0;
    }
    __syncthreads();
}

// iteration.c:250-293 (general variant)
template <bool DIAG, bool CHECK, class WG>
__device__ __forceinline__ void wg_step_t(WG &g, double h1, double h2, double *sums, double *maxs)
{
    const State &S = *g.S;
    const int m = g.I->m;
    const double sg = S.sigma;
    const double f2s = 1.0 / (1.0 + sg * S.lam_q), f1s = 1.0 - f2s;
    wg_product(g, g.Q, g.N(N_w), g.I->n, EpiZxw1<DIAG, CHECK>{{}, g.N(N_Qw), g.N(N_x), g.N(N_x_bar), g.N(N_x_hat), g.N(N_w_bar), g.N(N_dx),
                                 g.N(N_z_bar), g.N(N_w), g.N(N_ATy), g.N(N_c), g.N(N_l), g.N(N_u), g.N(N_last_x),
                                 g.N(N_fact1), g.N(N_fact2), sg, f1s, f2s, h1, h2},
            sums, maxs);
    if (m > 0) {
        wg_product(g, g.Q, g.N(N_w_bar), g.I->n, EpiTempv{{}, g.N(N_tempv), g.N(N_x_hat), g.N(N_Qw), sg}, sums, maxs);
        const double f1 = S.lam_a * sg, f2 = 1.0 / f1;
        wg_product(g, g.A, g.N(N_tempv), g.I->n, EpiY<CHECK>{{}, g.M(M_y), g.M(M_y_bar), g.M(M_s), g.M(M_dy), g.M(M_AL), g.M(M_AU), g.M(M_last_y), f1,
                            f2, h1, h2},
                sums, maxs);
    }
    const double fs = sg / (1.0 + sg * S.lam_q);
    wg_product(g, g.AT, g.M(M_y_bar), m, EpiW2<DIAG, CHECK>{{}, g.N(N_w_bar), g.N(N_w), g.N(N_ATy), g.N(N_ATy_bar), g.N(N_dw), g.N(N_ATdy),
                               g.N(N_last_w), g.N(N_last_ATy), g.N(N_fact), fs, h1, h2},
            sums, maxs);
}

// iteration.c:19-83 and 176-216 (noC: A^T y first, Halpern on w and Qw)
template <bool DIAG, bool CHECK, class WG>
__device__ __forceinline__ void wg_step_noc_t(WG &g, double h1, double h2, double *sums, double *maxs)
{
    const State &S = *g.S;
    const int m = g.I->m;
    const double sg = S.sigma;
    const double f2s = 1.0 / (1.0 + sg * S.lam_q), f1s = 1.0 - f2s;
    wg_product(g, g.AT, g.M(M_y), m, EpiZxw1NoC<DIAG, CHECK>{{}, g.N(N_ATy), g.N(N_x), g.N(N_x_bar), g.N(N_x_hat), g.N(N_w), g.N(N_w_bar),
                                    g.N(N_dx), g.N(N_dw), g.N(N_z_bar), g.N(N_Qw), g.N(N_c), g.N(N_l), g.N(N_u),
                                    g.N(N_last_x), g.N(N_last_w), g.N(N_fact1), g.N(N_fact2), sg, f1s, f2s, h1, h2},
            sums, maxs);
    wg_product(g, g.Q, g.N(N_w_bar), g.I->n, EpiQwNoC{{}, g.N(N_Qw), g.N(N_Qw_bar), g.N(N_tempv), g.N(N_last_Qw), g.N(N_x_hat), sg, h1, h2}, sums,
            maxs);
    if (m > 0) {
        const double f1 = S.lam_a * sg, f2 = 1.0 / f1;
        wg_product(g, g.A, g.N(N_tempv), g.I->n, EpiY<CHECK>{{}, g.M(M_y), g.M(M_y_bar), g.M(M_s), g.M(M_dy), g.M(M_AL), g.M(M_AU), g.M(M_last_y), f1,
                            f2, h1, h2},
                sums, maxs);
    }
}

template <class WG>
__device__ __forceinline__ void wg_step(WG &g, long long inner, double *sums, double *maxs)
{
    const double h1 = 1.0 / ((double)inner + 2.0), h2 = 1.0 - h1;
    const bool diag = g.S->Q_is_diag, chk = g.S->to_check;
    if (g.I->noC) {
        if (diag) {
            if (chk) wg_step_noc_t<true, true>(g, h1, h2, sums, maxs);
            else wg_step_noc_t<true, false>(g, h1, h2, sums, maxs);
        } else {
            if (chk) wg_step_noc_t<false, true>(g, h1, h2, sums, maxs);
            else wg_step_noc_t<false, false>(g, h1, h2, sums, maxs);
        }
        return;
    }
    if (diag) {
        if (chk) wg_step_t<true, true>(g, h1, h2, sums, maxs);
        else wg_step_t<true, false>(g, h1, h2, sums, maxs);
    } else {
        if (chk) wg_step_t<false, true>(g, h1, h2, sums, maxs);
        else wg_step_t<false, false>(g, h1, h2, sums, maxs);
    }
}

// solver_backend.c:432-513, one body; returns true when the instance is done.
template <class WG>
__device__ __forceinline__ bool wg_body(WG &g, double *sums, double *maxs)
{
    State &S = *g.S;
    const long long iter = S.iter;
    if (iter > 0) {
        const bool pr = d_print_now(g.P, iter);
        if ((g.P.check_iter > 0 && iter % g.P.check_iter == 0) || pr) {
            wg_residuals(g, iter, sums, maxs);
            if (threadIdx.x == 0) {
                if (S.first4 && S.kkt < 1e-4) { S.iter4 = iter; S.first4 = 0; }
                if (S.first6 && S.kkt < 1e-6) { S.iter6 = iter; S.first6 = 0; }
            }
        }
    }
    if (threadIdx.x == 0) {
        int st = HPRQP_STATUS_CONTINUE;
        if (S.kkt < g.P.stoptol) st = HPRQP_STATUS_OPTIMAL;
        else if (iter == g.P.max_iter) st = HPRQP_STATUS_MAX_ITER;
        S.status = st;
        d_restart_check(S.rs, iter, g.P.check_iter, S.sigma);
    }
    __syncthreads();
    wg_update_sigma(g, sums, maxs);
    wg_restart_apply(g, sums, maxs);
    if (S.status != HPRQP_STATUS_CONTINUE) {
        __syncthreads();
        if (threadIdx.x == 0) S.finished = 1;
        __syncthreads();
        return true;
    }
    if (threadIdx.x == 0) S.to_check = d_check_next(g.P, S.rs, iter + 1);
    __syncthreads();
    wg_step(g, S.rs.inner, sums, maxs);
    if (S.rs.flag > 0 || (g.P.check_iter > 0 && (iter + 1) % g.P.check_iter == 0)) {
        const double gap = wg_m_norm(g, sums, maxs);
        if (threadIdx.x == 0) {
            if (!isfinite(gap)) {
                S.status = HPRQP_STATUS_INVALID_ARGUMENT;
                S.finished = 1;
            } else {
                d_restart_post(S.rs, iter, g.P.check_iter, gap);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        S.rs.inner += 1;
        S.iter = iter + 1;
    }
    __syncthreads();
    return S.finished != 0;
}

template <class WG>
__device__ __forceinline__ void wg_bind(WG &g, const Arrays &ar, const Params &P, int k, State *S, double *prod, double *red)
{
    const Inst *I = ar.inst + k;
    g.I = I;
    g.S = S;
    g.P = P;
    g.prod = prod;
    g.red = red;
    g.Q = Mat{ar.rp + I->q_rp, ar.ci + I->q_nz, ar.val + I->q_nz, ar.blocks + I->q_blk, I->q_nblk,
              nullptr, 0, 0, 0};
    g.A = Mat{ar.rp + I->a_rp, ar.ci + ar.a_base + I->a_nz, ar.val + ar.a_base + I->a_nz, ar.blocks + I->a_blk,
              I->a_nblk, nullptr, 0, 0, 0};
    g.AT = Mat{ar.rp + I->at_rp, ar.ci + ar.at_ci_base + I->a_nz, ar.val + ar.at_ci_base + I->a_nz,
               ar.blocks + I->at_blk, I->at_nblk, nullptr, 0, 0, 0};
    g.Q.wave_row = g.A.wave_row = g.AT.wave_row = ar.wave_row;
    if constexpr (WG::kDense) {
        // column form: Q over Q (symmetric), A x over A^T (n x m), A^T y over A (m x n)
        g.Q.dv = ar.dval + I->dq;
        g.Q.K = I->n, g.Q.J = I->n, g.Q.ld = I->n;
        g.A.dv = ar.dval + I->da;
        g.A.K = I->n, g.A.J = I->m, g.A.ld = I->m;
        g.AT.dv = ar.dval + I->dat;
        g.AT.K = I->m, g.AT.J = I->n, g.AT.ld = I->n;
        g.G = Mat{nullptr, nullptr, nullptr, nullptr, 0, ar.gram ? ar.gram + I->dg : nullptr, I->m, I->m, I->m, 0};
    }
    g.a_val = ar.val + ar.a_base + I->a_nz;
    g.at_perm = ar.perm + I->a_nz;
    g.qt_rp = ar.rp + I->qt_rp;
    g.qt_perm = ar.perm + ar.qt_perm_base + I->q_nz;
    for (int v = 0; v < N_COUNT; ++v) g.vn[v] = ar.vn + (long long)v * ar.sum_n + I->vn;
    for (int v = 0; v < M_COUNT; ++v) g.vm[v] = ar.vm + (long long)v * ar.sum_m + I->vm;
}


// ------------------------------------------------------------- adjoint
// Implicit differentiation of the converged KKT point of one instance.
// Active set from the projections: row i is active when y_bar_i != 0 or it is
// an equality row; variable j is bound-active when z_bar_j != 0 or l_j = u_j.
// In the scaled space (Q~ = D^-1 Q D^-1, A~ = R^-1 A D^-1) solve
//   [Q~_FF  A~_SF'] [u~_F]   [(g/d)_F]
//   [A~_SF    0   ] [v~_S] = [   0   ]
// matrix-free by MINRES (the conjugate-residual form, one operator application
// per step) with the workgroup SpMV engine, then recover the bound
// multipliers v~_B = (g/d)_B - (Q~ u~ + A~' v~)_B and unscale:
// u = u~/d, v_rows = v~/r, v_bounds = d v~_B.

struct AdjMask {
    const double *z_bar, *l, *u, *y_bar, *AL, *AU;
    __device__ bool free_var(int j) const { return z_bar[j] == 0.0 && l[j] != u[j]; }
    __device__ bool active_row(int i) const { return y_bar[i] != 0.0 || AL[i] == AU[i]; }
};

struct EpiAdjN : NoRed {  // yn = mask_F (tmp + A~' xm)
    double *yn;
    const double *tmp;
    AdjMask mk;
    __device__ void row(int i, double v, double *, double *) const { yn[i] = mk.free_var(i) ? tmp[i] + v : 0.0; }
};

struct EpiAddStore : NoRed {  // out = tmp + value
    double *out;
    const double *tmp;
    __device__ void row(int i, double v, double *, double *) const { out[i] = tmp[i] + v; }
};

struct EpiAdjM : NoRed {  // ym = mask_S (A~ xn)
    double *ym;
    AdjMask mk;
    __device__ void row(int i, double v, double *, double *) const { ym[i] = mk.active_row(i) ? v : 0.0; }
};

template <class WG>
__device__ void adj_op(WG &g, const AdjMask &mk, const double *xn, const double *xm, double *yn, double *ym,
                       double *sums, double *maxs)
{
    double *tmp = g.N(N_pt);
    wg_spmv<WG::kDense>(g.Q, xn, EpiStore{{}, tmp}, sums, maxs, g.prod);
    wg_spmv<WG::kDense>(g.AT, xm, EpiAdjN{{}, yn, tmp, mk}, sums, maxs, g.prod);
    if (g.I->m > 0) wg_spmv<WG::kDense>(g.A, xn, EpiAdjM{{}, ym, mk}, sums, maxs, g.prod);
}

// Sums of kAdjRows vectors over a wavefront in one pass (a reduce-scatter
// by halves, then exchanges over the remaining low lane bits): v[u] holds
// lane's term of row u on entry; on exit lane L holds the wavefront's sum
// of row L / (kWave / kAdjRows) in v[0]. Fixed pairing, so deterministic.
constexpr int kAdjRows = 16;  // rows of A per chunk: loads in flight per thread
constexpr int kAdjStages = 4;  // log2(kAdjRows)
static_assert((1 << kAdjStages) == kAdjRows && kAdjRows <= kWave, "reduce-scatter shape");

__device__ inline void wave_reduce_scatter(double (&v)[kAdjRows], int lane)
{
#pragma unroll
    for (int st = 0; st < kAdjStages; ++st) {
        const int o = (kWave / 2) >> st;
        const int h = kAdjRows >> (st + 1);
        const bool up = (lane & o) != 0;
#pragma unroll
        for (int k = 0; k < h; ++k) {
            const double mine = up ? v[h + k] : v[k];
            const double send = up ? v[k] : v[h + k];
            v[k] = mine + __shfl_xor(send, o, kWave);
        }
    }
#pragma unroll
    for (int o = (kWave / kAdjRows) / 2; o > 0; o >>= 1) v[0] += __shfl_xor(v[0], o, kWave);
}

// The same operator on a dense instance in one pass over A instead of two
// (Q: n x n, A: m x n, row-major; g.AT's value plane is A itself): thread j
// of a 256-column tile adds A[i][j] xm_i to its column sum, i ascending (the
// order of the column-form A' product), and the row products A[i][j] xn_j
// are summed over the workgroup -- each wavefront's 64 columns by
// wave_reduce_scatter into wrow[wavefront][i], then the four wavefronts in
// order, then the column tiles in order -- into ym_i. Q xn stays a
// column-form product. The next chunk's rows of A are loaded before the
// current chunk is reduced, and no barrier interrupts the pass. x_n and x_m
// are staged in LDS (xs_n, xs_m: kDenseMax each); wrow (4 x mld doubles,
// dynamic LDS) holds the wavefronts' partial row sums.
template <class WG>
__device__ void adj_op_dense(WG &g, const AdjMask &mk, const double *xn, const double *xm, double *yn, double *ym,
                             double *xs_n, double *xs_m, double *wrow, int mld)
{
    const int n = g.I->n, m = g.I->m, t = threadIdx.x;
    const int lane = t & (kWave - 1), wid = t / kWave;
    for (int i = t; i < n; i += kThreads) xs_n[i] = xn[i];
    for (int i = t; i < m; i += kThreads) xs_m[i] = xm[i];
    __syncthreads();
    const double *__restrict__ Ad = g.AT.dv;  // A[i][j] at i * n + j
    double *__restrict__ wr = wrow + wid * mld;
    for (int j0 = 0; j0 < n; j0 += kThreads) {
        const int j = j0 + t;
        const bool col = j < n;
        const int jc = col ? j : n - 1;
        const double xj = col ? xs_n[j] : 0.0;
        const double qx = dense_column_dot<false>(g.Q, jc, xs_n);
        double acc = 0.0;
        double a[kAdjRows];
        if (m > 0) {
#pragma unroll
            for (int u = 0; u < kAdjRows; ++u)
                a[u] = __builtin_nontemporal_load(Ad + (long long)min(u, m - 1) * n + jc);
        }
        for (int i0 = 0; i0 < m; i0 += kAdjRows) {
            double p[kAdjRows];
#pragma unroll
            for (int u = 0; u < kAdjRows; ++u) {
                const bool ok = i0 + u < m;
                if (ok) acc += a[u] * xs_m[i0 + u];
                p[u] = col && ok ? a[u] * xj : 0.0;
            }
            if (i0 + kAdjRows < m) {  // next chunk in flight during the reduction
#pragma unroll
                for (int u = 0; u < kAdjRows; ++u)
                    a[u] = __builtin_nontemporal_load(Ad + (long long)min(i0 + kAdjRows + u, m - 1) * n + jc);
            }
            wave_reduce_scatter(p, lane);
            constexpr int kLanesPerRow = kWave / kAdjRows;
            const int r = i0 + lane / kLanesPerRow;
            if (lane % kLanesPerRow == 0 && r < m) wr[r] = p[0];
        }
        if (col) yn[j] = mk.free_var(j) ? qx + acc : 0.0;
        __syncthreads();
        for (int i = t; i < m; i += kThreads) {
            double r = wrow[i];
#pragma unroll
            for (int w = 1; w < kThreads / kWave; ++w) r += wrow[w * mld + i];
            ym[i] = j0 == 0 ? r : ym[i] + r;
        }
        __syncthreads();
    }
    for (int i = t; i < m; i += kThreads)
        if (!mk.active_row(i)) ym[i] = 0.0;
    __syncthreads();
}

template <class WG>
__device__ double adj_dot(WG &g, const double *an, const double *am, const double *bn, const double *bm,
                          double *sums, double *maxs)
{
    const int n = g.I->n, m = g.I->m;
    for (int i = threadIdx.x; i < n; i += kThreads) sums[0] += an[i] * bn[i];
    for (int i = threadIdx.x; i < m; i += kThreads) sums[0] += am[i] * bm[i];
    __syncthreads();
    wg_reduce<1, 0>(sums, maxs, g.red);
    const double v = g.red[0];
    __syncthreads();
    return v;
}

// Sum over the 256-thread workgroup (DPP wavefront sums, then the four
// wavefronts in order); valid in every thread. red: kThreads / kWave doubles
// used by nothing else.
__device__ double adj_block_sum(double v, double *red)
{
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    v = wave_sum_dpp(v);
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = red[0];
#pragma unroll
    for (int w = 1; w < kThreads / kWave; ++w) t += red[w];
    __syncthreads();  // red is rewritten by the next sum
    return t;
}

template <bool D>
__global__ __launch_bounds__(kThreads) void batch_adjoint_kernel(Arrays ar, Params P, int count,
                                                                 const double *__restrict__ gvec,
                                                                 double *__restrict__ u_out,
                                                                 double *__restrict__ vrow_out,
                                                                 double *__restrict__ vbnd_out, double tol,
                                                                 int max_steps, int *__restrict__ steps_out,
                                                                 double *__restrict__ resid_out, int mld)
{
    __shared__ double prod[kStreamCap];
    __shared__ double red[kMaxTerms];
    __shared__ double adj_red[kThreads / kWave];
    __shared__ State S;
    // dense instances: x_m staging of the one-pass operator (adj_op_dense)
    __shared__ double dense_lds[D ? kDenseMax : 1];
    extern __shared__ double adj_dyn[];  // dense: 4 x mld partial row sums
    if ((int)blockIdx.x >= count) return;
    const int k = ar.order[blockIdx.x];
    if (ar.inst[k].dense != (int)D) return;
    if (threadIdx.x == 0) S = ar.state[k];
    __syncthreads();
    WGv<D> g;
    wg_bind(g, ar, P, k, &S, prod, red);
    const int n = g.I->n, m = g.I->m, t = threadIdx.x;
    double sums[kMaxTerms] = {0}, maxs[kMaxTerms] = {0};
    auto apply = [&](const AdjMask &mk_, const double *xn_, const double *xm_, double *yn_, double *ym_) {
        if constexpr (D)
            adj_op_dense(g, mk_, xn_, xm_, yn_, ym_, prod, dense_lds, adj_dyn, mld);
        else
            adj_op(g, mk_, xn_, xm_, yn_, ym_, sums, maxs);
    };
    const AdjMask mk{g.N(N_z_bar), g.N(N_l), g.N(N_u), g.M(M_y_bar), g.M(M_AL), g.M(M_AU)};
    const double *d = g.N(N_col), *rr = g.M(M_row);
    const double *gk = gvec + g.I->vn;
    double *bn = g.N(N_tc);
    // MINRES (Paige and Saunders, SIAM J. Numer. Anal. 12 (1975) 617-629):
    // Lanczos on the symmetric indefinite active-set operator with Givens
    // rotations of the tridiagonal factor, so each step takes one operator
    // product, the Lanczos vectors stay normalised and the iterate minimises
    // the residual over the Krylov space. Vectors (free part, row part):
    // x; v (normalised Lanczos vector); r1, r2, y (Lanczos three-term
    // recurrence, rotated); w, w1, w2 (update directions, rotated).
    double *xn = g.N(N_pq), *xm = g.M(M_pq);
    double *vn = g.N(N_pz), *vm = g.M(M_pz);
    // rotated by swapping pointers (no indexed pointer arrays: those would
    // live in scratch memory)
    double *r1n = g.N(N_dx), *r2n = g.N(N_dw), *yn = g.N(N_tempv);
    double *r1m = g.M(M_tr), *r2m = g.M(M_dy), *ym = g.M(M_e0);
    double *wn = g.N(N_dQw), *w1n = g.N(N_QATdy), *w2n = g.N(N_Qx);
    double *wm = g.M(M_e1), *w1m = g.M(M_e2), *w2m = g.M(M_e3);
    for (int i = t; i < n; i += kThreads) {
        const double b = mk.free_var(i) ? gk[i] / d[i] : 0.0;
        bn[i] = b;
        xn[i] = 0.0;
        r1n[i] = b;
        r2n[i] = b;
        wn[i] = 0.0;
        w2n[i] = 0.0;
    }
    for (int i = t; i < m; i += kThreads) {
        xm[i] = 0.0;
        r1m[i] = 0.0;
        r2m[i] = 0.0;
        wm[i] = 0.0;
        w2m[i] = 0.0;
    }
    __syncthreads();
    const double beta1 = sqrt(adj_dot(g, bn, r1m, bn, r1m, sums, maxs));
    double oldb = 0.0, beta = beta1, dbar = 0.0, epsln = 0.0, phibar = beta1, cs = -1.0, sn = 0.0;
    int step = 0;
    if (beta1 > 0.0) {
        // Each thread owns the same indices in every elementwise pass, so a
        // pass needs no barrier before the next one reads what it wrote;
        // barriers stand only before the operator (it reads v at every
        // index) and inside the two sums. A step is then: the operator,
        // (y -= (beta/oldb) r1, alfa = v.y), (y -= (alfa/beta) r2,
        // beta = |y|), (w and x updates, next v = y/beta).
        {
            const double sc = 1.0 / beta;
            for (int i = t; i < n; i += kThreads) vn[i] = sc * r2n[i];
            for (int i = t; i < m; i += kThreads) vm[i] = sc * r2m[i];
        }
        __syncthreads();
        for (step = 1; step <= max_steps; ++step) {
            apply(mk, vn, vm, yn, ym);
            __syncthreads();
            double pa = 0.0;
            {
                const double a1 = step >= 2 ? -beta / oldb : 0.0;
                for (int i = t; i < n; i += kThreads) {
                    const double y = step >= 2 ? yn[i] + a1 * r1n[i] : yn[i];
                    yn[i] = y;
                    pa += vn[i] * y;
                }
                for (int i = t; i < m; i += kThreads) {
                    const double y = step >= 2 ? ym[i] + a1 * r1m[i] : ym[i];
                    ym[i] = y;
                    pa += vm[i] * y;
                }
            }
            const double alfa = adj_block_sum(pa, adj_red);
            double pb = 0.0;
            {
                const double a2 = -alfa / beta;
                for (int i = t; i < n; i += kThreads) {
                    const double y = yn[i] + a2 * r2n[i];
                    yn[i] = y;
                    pb += y * y;
                }
                for (int i = t; i < m; i += kThreads) {
                    const double y = ym[i] + a2 * r2m[i];
                    ym[i] = y;
                    pb += y * y;
                }
            }
            {  // r1 <- r2, r2 <- y; the next y reuses r1's buffer
                double *on = r1n, *om = r1m;
                r1n = r2n, r1m = r2m;
                r2n = yn, r2m = ym;
                yn = on, ym = om;
            }
            oldb = beta;
            beta = sqrt(adj_block_sum(pb, adj_red));
            const double oldeps = epsln;
            const double delta = cs * dbar + sn * alfa;
            const double gbar = sn * dbar - cs * alfa;
            epsln = sn * beta;
            dbar = -cs * beta;
            const double gamma = fmax(sqrt(gbar * gbar + beta * beta), 1e-300);
            cs = gbar / gamma;
            sn = beta / gamma;
            const double phi = cs * phibar;
            phibar = sn * phibar;
            {  // w1 <- w2, w2 <- w; the new w goes into w1's old buffer
                double *on = w1n, *om = w1m;
                w1n = w2n, w1m = w2m;
                w2n = wn, w2m = wm;
                wn = on, wm = om;
            }
            const double rg = 1.0 / gamma, sc = beta > 0.0 ? 1.0 / beta : 0.0;
            for (int i = t; i < n; i += kThreads) {
                const double w = (vn[i] - oldeps * w1n[i] - delta * w2n[i]) * rg;
                wn[i] = w;
                xn[i] += phi * w;
                vn[i] = sc * r2n[i];
            }
            for (int i = t; i < m; i += kThreads) {
                const double w = (vm[i] - oldeps * w1m[i] - delta * w2m[i]) * rg;
                wm[i] = w;
                xm[i] += phi * w;
                vm[i] = sc * r2m[i];
            }
            __syncthreads();
            if (phibar <= tol * beta1 || !(beta > 0.0)) break;  // converged, or the Krylov space is exhausted
        }
        step = min(step, max_steps);
    }
    // the true relative residual ||b - K x|| / ||b|| of the returned iterate
    double rel = 0.0;
    if (beta1 > 0.0) {
        apply(mk, xn, xm, yn, ym);
        __syncthreads();
        for (int i = t; i < n; i += kThreads) yn[i] = bn[i] - yn[i];
        for (int i = t; i < m; i += kThreads) ym[i] = -ym[i];
        __syncthreads();
        rel = sqrt(adj_dot(g, yn, ym, yn, ym, sums, maxs)) / beta1;
    }
    // bound multipliers: v~_B = b~_B - (Q~ u~ + A~' v~)_B with b~ = g/d
    double *tn = g.N(N_Qw_bar), *tmp = g.N(N_pt);
    wg_spmv<D>(g.Q, xn, EpiStore{{}, tmp}, sums, maxs, g.prod);
    wg_spmv<D>(g.AT, xm, EpiAddStore{{}, tn, tmp}, sums, maxs, g.prod);
    for (int i = t; i < n; i += kThreads) {
        const bool fr = mk.free_var(i);
        u_out[g.I->vn + i] = xn[i] / d[i];
        vbnd_out[g.I->vn + i] = fr ? 0.0 : d[i] * (gk[i] / d[i] - tn[i]);
    }
    for (int i = t; i < m; i += kThreads) vrow_out[g.I->vm + i] = xm[i] / rr[i];
    if (t == 0) {
        steps_out[k] = step;
        resid_out[k] = rel;
    }
}

// ------------------------------------------------ power iterations (CSR)
// eigs.c:19-153 for CSR instances with n, m <= kPowMax, one 512-thread
// workgroup per instance (256 VGPRs a thread), vectors in LDS. When a matrix
// has at most kPowCache entries per thread, thread t holds the contiguous
// run of entries [t C, (t + 1) C) (C = ceil(nnz / 512), row-major order) in
// registers for the whole iteration, so a power step reads no memory: each
// thread sums the rows of its run left to right (product, then add: the
// rounding of the host's CSR product without a fused multiply-add) and
// writes the rows that lie wholly in its run; a row split over several runs
// is finished by the thread where it starts, adding the pieces the later
// runs left in LDS in order. Larger matrices take the thread-per-row
// product from memory.
constexpr int kPowThreads = 512;
constexpr int kPowMax = 4096;    // longest vector of an instance the kernel takes
constexpr int kPowSlots = kPowMax / kPowThreads;  // rows a thread owns: t + kPowThreads s
constexpr int kPowCache = 24;    // matrix entries per thread held in registers
constexpr int kPowGather = 8;    // gathers a thread issues together
static_assert(kPowCache % kPowGather == 0, "runs are walked in whole gather groups");
constexpr int kPowRed = 4 * (kPowThreads / kWave);
// doubles of LDS: q, z and A' q (n, m <= kPowMax), the run pieces and two
// sets of mid flags
constexpr int kPowLds = 3 * kPowMax + 2 * kPowThreads;

// Both workgroup sums of (a, b) in one pass; slots 2 * slot and 2 * slot + 1.
__device__ double2 pow_block_sum2(double a, double b, double *red, int slot)
{
    constexpr int W = kPowThreads / kWave;
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    a = wave_sum_dpp(a);
    b = wave_sum_dpp(b);
    if (lane == 0) {
        red[2 * slot * W + wid] = a;
        red[(2 * slot + 1) * W + wid] = b;
    }
    __syncthreads();
    double2 t{red[2 * slot * W], red[(2 * slot + 1) * W]};
#pragma unroll
    for (int w = 1; w < W; ++w) {
        t.x += red[2 * slot * W + w];
        t.y += red[(2 * slot + 1) * W + w];
    }
    return t;
}

__device__ double pow_block_sum(double v, double *red, int slot)
{
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    v = wave_sum_dpp(v);
    if (lane == 0) red[slot * (kPowThreads / kWave) + wid] = v;
    __syncthreads();
    double t = red[slot * (kPowThreads / kWave)];
#pragma unroll
    for (int w = 1; w < kPowThreads / kWave; ++w) t += red[slot * (kPowThreads / kWave) + w];
    return t;
}

// y = M x over nrow rows of a CSR instance view, x in LDS; f(r, y_r).
template <class F>
__device__ void pow_spmv(const Mat &M, int nrow, const double *xs, const F &f)
{
    for (int r = threadIdx.x; r < nrow; r += kPowThreads) {
        const int k0 = M.rp[r], k1 = M.rp[r + 1];
        double acc = 0.0;
        int k = k0;
        for (; k + 8 <= k1; k += 8) {
            int c[8];
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                c[u] = M.ci[k + u];
                v[u] = M.val[k + u];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u] * xs[c[u]];
        }
        for (; k < k1; ++k) acc += M.val[k] * xs[M.ci[k]];
        f(r, acc);
    }
}

// One thread's run of a CSR view (R rows, Z entries), held in registers:
// entries [t C, t C + cnt) with cnt = min(C, Z - t C).
struct PowRun {
    double v[kPowCache];
    int rc[kPowCache];  // row | col << 16 (R, columns <= 4096)
    // first row | last row << 12 | head << 24 | tail << 25: rows of the
    // first and last held entry; head: the first row began in an earlier
    // run; tail: the last row goes on past the run
    unsigned meta;
    unsigned empty;  // bit s: row t + kPowThreads s has no entries

    __device__ int first() const { return meta & 0xfff; }
    __device__ int last() const { return (meta >> 12) & 0xfff; }
    __device__ bool head() const { return meta >> 24 & 1u; }
    __device__ bool tail() const { return meta >> 25 & 1u; }
    __device__ static int count(int Z, int C)
    {
        const int left = Z - (int)threadIdx.x * C;
        return left < C ? (left > 0 ? left : 0) : C;
    }

    // mid[t] (LDS): the run lies inside one row that goes on past it.
    __device__ void load(const Mat &M, int R, int Z, int C, int *mid)
    {
        const int t = threadIdx.x, lo = t * C, cnt = count(Z, C);
        unsigned empty = 0;
#pragma unroll
        for (int s = 0; s < kPowSlots; ++s) {
            const int r = t + s * kPowThreads;
            if (r < R && M.rp[r] == M.rp[r + 1]) empty |= 1u << s;
        }
        int row = 0;
        if (cnt > 0) {
            int hi = R;  // M.rp[row] <= lo < M.rp[hi]
            while (hi - row > 1) {
                const int mid_r = (row + hi) >> 1;
                if (M.rp[mid_r] <= lo) row = mid_r;
                else hi = mid_r;
            }
        }
        const int first_row = row;
#pragma unroll
        for (int u = 0; u < kPowCache; ++u) {
            v[u] = 0.0;
            rc[u] = 0;
            if (u < cnt) {
                const int e = lo + u;
                while (M.rp[row + 1] <= e) ++row;
                v[u] = M.val[e];
                rc[u] = row | (M.ci[e] << 16);
            }
        }
        const bool head = cnt > 0 && M.rp[first_row] < lo, tail = cnt > 0 && M.rp[row + 1] > lo + cnt;
        meta = (unsigned)first_row | (unsigned)row << 12 | (unsigned)head << 24 | (unsigned)tail << 25;
        this->empty = empty;
        mid[t] = head && tail && first_row == row;
    }

    // y = M x for the rows this run finishes; pieces of split rows go to
    // H[t] (a row begun in an earlier run) or `own` (the row this run
    // begins). combine() finishes the split rows after a barrier.
    __device__ void run(const double *__restrict__ x, double *__restrict__ y, double *__restrict__ H, double &own,
                        int Z, int C) const
    {
        const int t = threadIdx.x, cnt = count(Z, C);
#pragma unroll
        for (int s = 0; s < kPowSlots; ++s)
            if (empty >> s & 1u) y[t + s * kPowThreads] = 0.0;
        if (cnt == 0) return;
        double acc = 0.0;
        int cur = first();
#pragma unroll
        for (int u0 = 0; u0 < kPowCache; u0 += kPowGather) {
            // the gathers of a group go out together (an entry past the run
            // reads x[0] and is not used); the empty asm keeps the compiler
            // from holding every gather address in a register across steps
            double xv[kPowGather];
#pragma unroll
            for (int j = 0; j < kPowGather; ++j) {
                int rcu = rc[u0 + j];
                asm volatile("" : "+v"(rcu));
                xv[j] = x[u0 + j < cnt ? rcu >> 16 : 0];
            }
#pragma unroll
            for (int j = 0; j < kPowGather; ++j) {
                const int u = u0 + j;
                if (u < cnt) {
                    const int r = rc[u] & 0xffff;
                    if (r != cur) {  // cur is complete; only the first row can be a split piece here
                        if (head() && cur == first()) H[t] = acc;
                        else y[cur] = acc;
                        acc = 0.0;
                        cur = r;
                    }
                    const double p = v[u] * xv[j];
                    acc += p;
                }
            }
        }
        if (head() && cur == first()) H[t] = acc;
        else if (tail()) own = acc;
        else y[cur] = acc;
    }

    __device__ void combine(double *y, const double *H, const int *mid, double own, int Z, int C) const
    {
        if (count(Z, C) == 0 || !tail() || (head() && first() == last())) return;
        double total = own;
        for (int j = threadIdx.x + 1;; ++j) {
            total += H[j];
            if (!mid[j]) break;
        }
        y[last()] = total;
    }
};

// The power iteration of wg_power over vectors in LDS; prod() writes z = M q
// (q complete in LDS).
template <class Prod>
__device__ double pow_loop(int len, double *q, double *z, double *red, const Prod &prod)
{
    const int t = threadIdx.x;
    double s = 0.0;
    const double len1 = (double)len + 1.0;
    for (int i = t; i < len; i += kPowThreads) {
        const double v = 1.0 + ((double)i + 1.0) / len1;
        z[i] = v;
        s += v * v;
    }
    double qn = sqrt(pow_block_sum(s, red, 0));
    double lam = 1.0;
    for (int it = 0; it < 5000; ++it) {
        if (qn <= 1e-16) {
            lam = 0.0;
            break;
        }
        for (int i = t; i < len; i += kPowThreads) q[i] = z[i] / qn;
        __syncthreads();
        prod();
        __syncthreads();
        double s0 = 0.0, s1 = 0.0;
        for (int i = t; i < len; i += kPowThreads) {
            const double v = z[i];
            s0 += q[i] * v;
            s1 += v * v;
        }
        const double2 ss = pow_block_sum2(s0, s1, red, 1);  // slots 2 and 3
        lam = ss.x;
        const double zn = sqrt(ss.y);
        double r2 = 0.0;
        for (int i = t; i < len; i += kPowThreads) {
            const double r = z[i] - lam * q[i];
            q[i] = r;
            r2 += r * r;
        }
        const double err = sqrt(pow_block_sum(r2, red, 1)) / (zn + fabs(lam) + 1e-16);
        __syncthreads();  // red, q, z and the products are rewritten by the next step
        if (err < 1e-4) break;
        qn = zn;
    }
    return lam < 0.0 ? 0.0 : lam;
}

// lambda_max(A A'): z = A (A' q) over vectors of length m.
__device__ double pow_a(const Mat &A, const Mat &AT, int n, int m, double *lds, double *red)
{
    if (m == 0) return 0.0;
    const int Z = A.rp[m];
    double *q = lds, *z = lds + m, *tmp = lds + 2 * m;
    auto store = [](double *y) { return [y](int r, double v) { y[r] = v; }; };
    if (Z > kPowThreads * kPowCache)
        return pow_loop(m, q, z, red, [&] {
            pow_spmv(AT, n, q, store(tmp));
            __syncthreads();
            pow_spmv(A, m, tmp, store(z));
        });
    // q | z | tmp | run pieces | mid flags of A' | mid flags of A
    double *H = tmp + n;
    int *mid_at = reinterpret_cast<int *>(H + kPowThreads), *mid_a = mid_at + kPowThreads;
    const int C = (Z + kPowThreads - 1) / kPowThreads;
    PowRun at, a;
    at.load(AT, n, Z, C, mid_at);
    a.load(A, m, Z, C, mid_a);
    __syncthreads();
    return pow_loop(m, q, z, red, [&] {
        double own = 0.0;
        at.run(q, tmp, H, own, Z, C);
        __syncthreads();
        at.combine(tmp, H, mid_at, own, Z, C);
        __syncthreads();
        a.run(tmp, z, H, own, Z, C);
        __syncthreads();
        a.combine(z, H, mid_a, own, Z, C);
    });
}

// lambda_max(Q): z = Q q over vectors of length n.
__device__ double pow_q(const Mat &Q, int n, double *lds, double *red)
{
    const int Z = Q.rp[n];
    double *q = lds, *z = lds + n;
    if (Z > kPowThreads * kPowCache)
        return pow_loop(n, q, z, red, [&] { pow_spmv(Q, n, q, [z](int r, double v) { z[r] = v; }); });
    double *H = z + n;
    int *mid = reinterpret_cast<int *>(H + kPowThreads);
    const int C = (Z + kPowThreads - 1) / kPowThreads;
    PowRun run;
    run.load(Q, n, Z, C, mid);
    __syncthreads();
    return pow_loop(n, q, z, red, [&] {
        double own = 0.0;
        run.run(q, z, H, own, Z, C);
        __syncthreads();
        run.combine(z, H, mid, own, Z, C);
    });
}

// lambda_max(A A') and, for a non-diagonal Q, lambda_max(Q) of every CSR
// instance, x 1.05, into its state; runs between setup phases 1 and 3 (the
// matrices are scaled and A' holds its values).
__global__ __launch_bounds__(kPowThreads) void batch_power_kernel(Arrays ar, int count)
{
    __shared__ double lds[kPowLds];
    __shared__ double red[kPowRed];
    if ((int)blockIdx.x >= count) return;
    const int k = ar.order[blockIdx.x];
    const Inst *I = ar.inst + k;
    if (I->dense) return;
    const Mat Q{ar.rp + I->q_rp, ar.ci + I->q_nz, ar.val + I->q_nz, nullptr, 0, nullptr, 0, 0, 0};
    const Mat A{ar.rp + I->a_rp, ar.ci + ar.a_base + I->a_nz, ar.val + ar.a_base + I->a_nz, nullptr, 0, nullptr, 0, 0, 0};
    const Mat AT{ar.rp + I->at_rp, ar.ci + ar.at_ci_base + I->a_nz, ar.val + ar.at_ci_base + I->a_nz, nullptr, 0,
                 nullptr, 0, 0, 0};
    const int diag = ar.state[k].Q_is_diag;
    const double lam_a = pow_a(A, AT, I->n, I->m, lds, red) * 1.05;
    const double lam_q = diag ? 0.0 : pow_q(Q, I->n, lds, red) * 1.05;
    if (threadIdx.x == 0) {
        ar.state[k].lam_a = lam_a;
        ar.state[k].lam_q = lam_q;
    }
}

template <bool D>
// phase 0: the whole setup; 1: up to the power iterations (the state is
// stored); 2: from the power iterations on (the state is loaded); 3: as 2
// with the lambdas batch_power_kernel stored. Dense batches run 1,
// dense_gram_kernel, 2; CSR batches whose vectors fit in LDS run 1,
// batch_power_kernel, 3.
__global__ __launch_bounds__(kThreads) void batch_setup_kernel(Arrays ar, Params P, int count, int phase)
{
    __shared__ double prod[kStreamCap];
    __shared__ double red[kMaxTerms];
    __shared__ State S;
    if ((int)blockIdx.x >= count) return;
    const int k = ar.order[blockIdx.x];
    if (ar.inst[k].dense != (int)D) return;
    WGv<D> g;
    wg_bind(g, ar, P, k, &S, prod, red);
    if (threadIdx.x == 0) {
        if (phase >= 2) {
            S = ar.state[k];
        } else {
            S = State{};
            S.first4 = S.first6 = 1;
            S.kkt = INFINITY;
            S.rs.first = 1;
            S.rs.save_gap = S.rs.current_gap = S.rs.last_gap = S.rs.best_gap = INFINITY;
            S.rs.weighted_norm = INFINITY;
            S.rs.best_sigma = 1.0;
            S.status = HPRQP_STATUS_CONTINUE;
        }
    }
    __syncthreads();
    double sums[kMaxTerms] = {0}, maxs[kMaxTerms] = {0};
    if (phase < 2) wg_setup_pre(g, sums, maxs);
    if (phase == 1) {
        if (threadIdx.x == 0) ar.state[k] = S;
        return;
    }
    wg_setup_post(g, sums, maxs, phase == 3);
    // iteration-0 residuals (solver_backend.c:396)
    wg_residuals(g, 0, sums, maxs);
    if (threadIdx.x == 0) {
        if (S.kkt < 1e-4) { S.iter4 = 0; S.first4 = 0; }
        if (S.kkt < 1e-6) { S.iter6 = 0; S.first6 = 0; }
        S.rs.best_sigma = S.sigma;
        ar.state[k] = S;
    }
}

// Runs up to `bodies` loop bodies of every unfinished instance.
// Dense instances stream their matrices with many loads in flight, which
// needs more registers than four waves per SIMD leave; CSR instances are
// latency-bound (gathers, barriers) and run six waves per SIMD, which beats
// the spills that the smaller register budget costs.
template <bool D, bool ST>
__global__ __launch_bounds__(kThreads, D ? 2 : 6) void batch_solve_kernel(Arrays ar, Params P, int count, int bodies)
{
    // CSR: row-block products (and with ST the staged gathered vector: 25 KB
    // in all, six workgroups per CU); dense: x staged for the column products
    static_assert(!(D && ST), "dense products stage x in prod");
    __shared__ double prod[D ? kStreamCap : kBatchStreamCap];
    __shared__ double xs[ST ? kStageMax : 1];
    __shared__ double red[kMaxTerms];
    __shared__ State S;
    if ((int)blockIdx.x >= count) return;
    const int k = ar.order[blockIdx.x];
    if (ar.inst[k].dense != (int)D) return;
    if (threadIdx.x == 0) S = ar.state[k];
    __syncthreads();
    if (S.finished) return;
    WGv<D, ST> g;
    wg_bind(g, ar, P, k, &S, prod, red);
    g.xs = xs;
    double sums[kMaxTerms] = {0}, maxs[kMaxTerms] = {0};
    for (int b = 0; b < bodies; ++b)
        if (wg_body(g, sums, maxs)) break;
    if (threadIdx.x == 0) {
        ar.state[k] = S;
        if (!S.finished) atomicAdd(ar.active, 1);
    }
}

// The tail of a batch: once few unfinished instances are left (three per CU),
// each CSR instance runs on a workgroup of kLoneThreads threads, whose
// kLoneThreads / kThreads groups stage and sum different row blocks of a
// product at once (a lone instance's iteration is a chain of dependent block
// loads, not a stream). Rows are summed as in batch_solve_kernel; the
// reductions over the wider workgroup add their terms in another order. The
// solve path's device functions are forced inline: within the 128 VGPRs of a
// 1,024-thread workgroup the inliner would otherwise leave whole iteration
// steps as calls (0.9-1.4 kB of scratch per lane; 64 B inlined).
template <bool ST>
__global__ __launch_bounds__(kLoneThreads) void batch_solve_lone_kernel(Arrays ar, Params P, int count, int bodies)
{
    __shared__ double prod[(kLoneThreads / kThreads) * kBatchStreamCap];
    __shared__ double xs[ST ? kStageMax : 1];
    __shared__ double red[kMaxTerms];
    __shared__ State S;
    if ((int)blockIdx.x >= count) return;
    const int k = ar.order[blockIdx.x];
    if (ar.inst[k].dense) return;
    if (threadIdx.x == 0) S = ar.state[k];
    __syncthreads();
    if (S.finished) return;
    WGv<false, ST, kLoneThreads> g;
    wg_bind(g, ar, P, k, &S, prod, red);
    g.xs = xs;
    double sums[kMaxTerms] = {0}, maxs[kMaxTerms] = {0};
    for (int b = 0; b < bodies; ++b)
        if (wg_body(g, sums, maxs)) break;
    if (threadIdx.x == 0) {
        ar.state[k] = S;
        if (!S.finished) atomicAdd(ar.active, 1);
    }
}

// Timing probe for the CSR engine's wave threshold: the three products of an
// iteration (Q w, A x, A^T y) of every sparse instance, into scratch vectors.
__global__ __launch_bounds__(kThreads) void batch_probe_kernel(Arrays ar, Params P, int count)
{
    __shared__ double prod[kStreamCap];
    __shared__ double red[kMaxTerms];
    __shared__ State S;
    if ((int)blockIdx.x >= count) return;
    const int k = ar.order[blockIdx.x];
    if (ar.inst[k].dense) return;
    WGv<false> g;
    wg_bind(g, ar, P, k, &S, prod, red);
    double sums[kMaxTerms] = {0}, maxs[kMaxTerms] = {0};
    wg_spmv<false>(g.Q, g.N(N_w), EpiStore{{}, g.N(N_pt)}, sums, maxs, g.prod);
    if (g.I->m > 0) {
        wg_spmv<false>(g.A, g.N(N_x), EpiStore{{}, g.M(M_pz)}, sums, maxs, g.prod);
        wg_spmv<false>(g.AT, g.M(M_y), EpiStore{{}, g.N(N_pq)}, sums, maxs, g.prod);
    }
}

// Clears the unfinished-instance counter (a kernel, so that it can sit in a
// captured graph).
__global__ void batch_zero_counter_kernel(int *counter) { *counter = 0; }

// build.c:157-252 on dense row-major Q planes (one workgroup per instance
// row): Q_sym(r, c) = h(r, c) + h(c, r) with h = Q/2 dropped when |h| <=
// 1e-15, the sum zeroed when |sum| <= 1e-15. With at most two contributions
// per entry the order of the sum does not matter, so this equals the host
// build of the same matrix.
__global__ __launch_bounds__(kThreads) void dense_symmetrize_kernel(double *__restrict__ Qp, int n, int count)
{
    const long long row = blockIdx.x;
    if (row >= (long long)count * n) return;
    const long long k = row / n;
    const int r = (int)(row % n);
    double *Q = Qp + k * n * n;
    const double tol = 1e-15;
    for (int c = r + threadIdx.x; c < n; c += kThreads) {
        double h1 = 0.5 * Q[(long long)r * n + c], h2 = 0.5 * Q[(long long)c * n + r];
        if (fabs(h1) <= tol) h1 = 0.0;
        if (fabs(h2) <= tol) h2 = 0.0;
        double v = h1 + h2;
        if (fabs(v) <= tol) v = 0.0;
        Q[(long long)r * n + c] = v;
        Q[(long long)c * n + r] = v;
    }
}

// G = A A^T of every dense instance (after scaling, for its power
// iteration): one workgroup per instance, 32 x 32 output tiles of the upper
// triangle with both halves written, the inner dimension in LDS chunks of
// 32; each entry is summed over the columns in ascending order, so G is
// exactly symmetric.
__global__ __launch_bounds__(kThreads) void dense_gram_kernel(Arrays ar, int count)
{
    constexpr int T = 32;
    __shared__ double ta[T][T + 1], tb[T][T + 1];
    if ((int)blockIdx.x >= count) return;
    const Inst &I = ar.inst[ar.order[blockIdx.x]];
    if (!I.dense) return;
    const int n = I.n, m = I.m, t = threadIdx.x;
    const double *A = ar.dval + I.dat;  // row-major m x n
    double *G = ar.gram + I.dg;
    const int tr = t / (T / 4), tc = (t % (T / 4)) * 4;  // this thread's row and first of four columns
    const int tiles = (m + T - 1) / T;
    for (int bi = 0; bi < tiles; ++bi)
        for (int bj = bi; bj < tiles; ++bj) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            for (int j0 = 0; j0 < n; j0 += T) {
                for (int e = t; e < T * T; e += kThreads) {
                    const int r = e / T, c = e % T, j = j0 + c;
                    const int ra = bi * T + r, rb = bj * T + r;
                    ta[r][c] = (ra < m && j < n) ? A[(long long)ra * n + j] : 0.0;
                    tb[r][c] = (rb < m && j < n) ? A[(long long)rb * n + j] : 0.0;
                }
                __syncthreads();
                for (int c = 0; c < T; ++c) {
                    const double a = ta[tr][c];
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc[u] += a * tb[tc + u][c];
                }
                __syncthreads();
            }
            const int i = bi * T + tr;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i2 = bj * T + tc + u;
                if (i < m && i2 < m) {
                    G[(long long)i * m + i2] = acc[u];
                    G[(long long)i2 * m + i] = acc[u];
                }
            }
        }
}

// Per-instance facts of a dense instance that the host needs before setup:
// whether it fails build_from_QAbc's validation (src/shared/build.c:339-400,
// plus: no row of A without entries), its number of free variables (noC
// choice) and the largest finite |AL|, |AU| and |c| (original norms).
struct DenseStats {
    int bad;
    int free_count;
    double norm_b;
    double norm_c;
};

// The same facts computed on the device from device-resident dense arrays
// (one workgroup per instance; Q and A are read once, coalesced).
__global__ __launch_bounds__(kThreads) void dense_inspect_kernel(const double *__restrict__ Q, const double *__restrict__ c,
                                                                 const double *__restrict__ A,
                                                                 const double *__restrict__ AL,
                                                                 const double *__restrict__ AU,
                                                                 const double *__restrict__ l,
                                                                 const double *__restrict__ u, int n, int m,
                                                                 DenseStats *__restrict__ out)
{
    __shared__ int rownz[kDenseMax];
    __shared__ int s_bad, s_free;
    __shared__ double s_nb[kThreads], s_nc[kThreads];
    const long long k = blockIdx.x;
    const int t = threadIdx.x;
    for (int i = t; i < m; i += kThreads) rownz[i] = 0;
    if (t == 0) s_bad = s_free = 0;
    __syncthreads();
    int bad = 0, fr = 0;
    double nb = 0.0, nc = 0.0;
    const double *Qk = Q + k * n * n;
    for (int e = t; e < n * n; e += kThreads)
        if (!isfinite(Qk[e])) bad = 1;
    const double *Ak = A + k * m * n;
    for (int e = t; e < m * n; e += kThreads) {
        const double v = Ak[e];
        if (!isfinite(v)) bad = 1;
        if (v != 0.0) rownz[e / n] = 1;  // benign race: every writer stores 1
    }
    for (int i = t; i < m; i += kThreads) {
        const double al = AL[k * m + i], au = AU[k * m + i];
        if (isnan(al) || isnan(au) || al > au || (isinf(al) && al < 0 && isinf(au) && au > 0)) bad = 1;
        nb = fmax(nb, fmax(isinf(al) ? 0.0 : fabs(al), isinf(au) ? 0.0 : fabs(au)));
    }
    for (int j = t; j < n; j += kThreads) {
        const double cj = c[k * n + j], lj = l[k * n + j], uj = u[k * n + j];
        if (!isfinite(cj) || isnan(lj) || isnan(uj) || lj > uj) bad = 1;
        if (isinf(lj) && isinf(uj)) ++fr;
        nc = fmax(nc, fabs(cj));
    }
    __syncthreads();
    for (int i = t; i < m; i += kThreads)
        if (!rownz[i]) bad = 1;
    if (bad) atomicOr(&s_bad, 1);
    if (fr) atomicAdd(&s_free, fr);
    s_nb[t] = nb;
    s_nc[t] = nc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (t < w) {
            s_nb[t] = fmax(s_nb[t], s_nb[t + w]);
            s_nc[t] = fmax(s_nc[t], s_nc[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) out[k] = DenseStats{s_bad, s_free, s_nb[0], s_nc[0]};
}

// Original-unit solution into caller buffers: x = x_bar / col, z = z_bar *
// col (sum_n entries), y = y_bar / row (real_m entries).
__global__ __launch_bounds__(kThreads) void batch_unscale_kernel(const double *__restrict__ vn, const double *__restrict__ vm,
                                                                 long long sum_n, long long sum_m, long long real_m,
                                                                 double *__restrict__ x, double *__restrict__ y,
                                                                 double *__restrict__ z)
{
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i < sum_n) {
        const double col = vn[(long long)N_col * sum_n + i];
        if (x) x[i] = vn[(long long)N_x_bar * sum_n + i] / col;
        if (z) z[i] = vn[(long long)N_z_bar * sum_n + i] * col;
    }
    if (y && i < real_m) y[i] = vm[(long long)M_y_bar * sum_m + i] / vm[(long long)M_row * sum_m + i];
}

}  // namespace batch
}  // namespace hprqp

// ================================================================ host side

// Host staging arrays of batch creation: 2 MiB aligned and marked for
// transparent huge pages, so that their first touches fault once per 2 MiB
// instead of once per 4 KiB page. Not zero-filled.
struct HostFree {
    void operator()(void *p) const { std::free(p); }
};
template <class T>
using HostArray = std::unique_ptr<T[], HostFree>;

template <class T>
HostArray<T> host_array(size_t count)
{
    constexpr size_t kHuge = size_t(2) << 20;
    const size_t bytes = (sizeof(T) * std::max<size_t>(count, 1) + kHuge - 1) / kHuge * kHuge;
    void *p = std::aligned_alloc(kHuge, bytes);
    if (p) (void)madvise(p, bytes, MADV_HUGEPAGE);
    return HostArray<T>(static_cast<T *>(p));
}

struct HPRQPHipBatch {
    size_t count = 0;
    bool any_dense = false, any_sparse = false;  // which kernel variants have work
    bool stage_x = false;  // every CSR instance has n, m <= kStageMax: the solve stages gathers in LDS
    std::vector<double> work;                    // per-instance size estimate (dispatch order)
    int *d_order = nullptr;
    size_t blocks_cap = 0;                       // ints allocated at d_blocks
    hipGraphExec_t graph = nullptr;              // captured chunk (counter clear + solve launches)
    size_t graph_bodies = 0;
    hipGraphExec_t graph_lone = nullptr;         // the same with the tail's lone-instance kernel
    size_t graph_lone_bodies = 0;
    // unfinished instances at or below which the CSR instances continue in the
    // lone kernel: three per CU (two lone workgroups fit a CU at once; at
    // 4,096 instances of n <= 2,000 the lone kernel already beats the shared
    // one with 649 left: the solve 0.372 s with one per CU, 0.367 s with three)
    int lone_max = 0;
    hprqp::batch::Arrays ar{};
    hprqp::batch::Params P{};
    std::vector<hprqp::batch::Inst> inst;
    std::vector<size_t> n, m;
    hipStream_t stream = nullptr;
    double *d_vn = nullptr, *d_vm = nullptr, *d_val = nullptr, *d_dval = nullptr;
    int *d_rp = nullptr, *d_ci = nullptr, *d_perm = nullptr, *d_blocks = nullptr, *d_active = nullptr;
    int *d_steps = nullptr;        // adjoint step counts of the last adjoint
    double *d_adj_resid = nullptr;  // adjoint final relative residuals of the last adjoint
    double *d_gram = nullptr;  // G = A A^T of the dense instances (setup)
    mutable hipEvent_t caller_done = nullptr;  // last work queued on a caller's stream that reads this batch
    HostArray<int> h_rp;  // host copy of all CSR row pointers (block planning)
    // CSR instances' A row and column offsets in the block-diagonal matrix of
    // all of them (A' entries and the A', Q' permutations are formed on the
    // device at setup); empty when the host formed them
    std::vector<int2> tr_off;
    std::vector<int> max_row;  // longest CSR row of each instance (Q, A and A'); empty: unknown
    hprqp::batch::Inst *d_inst = nullptr;
    hprqp::batch::State *d_state = nullptr;
    std::vector<hprqp::batch::State> state;
    double time_limit = 3600.0;
    double solve_seconds = 0.0;
    size_t launches = 0;
    ~HPRQPHipBatch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (caller_done) {
            (void)hipEventSynchronize(caller_done);
            (void)hipEventDestroy(caller_done);
        }
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graph_lone) (void)hipGraphExecDestroy(graph_lone);
        if (d_order) (void)hipFree(d_order);
        for (void *p : {(void *)d_vn, (void *)d_vm, (void *)d_val, (void *)d_dval, (void *)d_rp, (void *)d_ci, (void *)d_perm,
                        (void *)d_blocks, (void *)d_active, (void *)d_inst, (void *)d_state, (void *)d_steps,
                        (void *)d_adj_resid, (void *)d_gram})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Row blocks for the workgroup engine: runs of rows with >= wave_row entries
// (up to 8 per block, one wavefront per row), otherwise stream blocks of
// <= kThreads rows and <= kBatchStreamCap entries.
std::vector<int> plan_batch_blocks(const int *rp, int nrow, int wave_row)
{
    const int kWaveRow = wave_row;
    std::vector<int> starts;
    int r = 0;
    while (r < nrow) {
        starts.push_back(r);
        const int first = r;
        if (rp[r + 1] - rp[r] >= kWaveRow) {
            while (r < nrow && r - first < 8 && rp[r + 1] - rp[r] >= kWaveRow) ++r;
            continue;
        }
        int nnz = 0;
        while (r < nrow && r - first < hprqp::kThreads) {
            const int len = rp[r + 1] - rp[r];
            if (len >= kWaveRow || nnz + len > hprqp::batch::kBatchStreamCap) break;
            nnz += len;
            ++r;
        }
    }
    starts.push_back(nrow);
    if (nrow == 0) starts = {0, 0};
    return starts;
}

// Row blocks of every CSR instance (Q, A, A^T) for wave threshold
// `wave_row`, planned on worker threads; sets the instances' block offsets
// and returns the block lists concatenated in instance order.
std::vector<int> plan_all_blocks(HPRQPHipBatch *b, const int *rp, int wave_row)
{
    const size_t count = b->count;
    std::vector<std::vector<int>> per(3 * count);
    {
        const size_t workers = std::max<size_t>(1, std::min<size_t>({16, count, (size_t)std::max(1u, std::thread::hardware_concurrency())}));
        std::atomic<size_t> next{0};
        auto work = [&] {
            for (size_t k; (k = next.fetch_add(1)) < count;) {
                const hprqp::batch::Inst &I = b->inst[k];
                if (I.dense) continue;
                per[3 * k] = plan_batch_blocks(rp + I.q_rp, I.n, wave_row);
                per[3 * k + 1] = plan_batch_blocks(rp + I.a_rp, I.m, wave_row);
                per[3 * k + 2] = plan_batch_blocks(rp + I.at_rp, I.n, wave_row);
            }
        };
        std::vector<std::thread> pool;
        for (size_t w = 1; w < workers; ++w) pool.emplace_back(work);
        work();
        for (std::thread &th : pool) th.join();
    }
    std::vector<int> blocks;
    size_t total = 0;
    for (const auto &v : per) total += v.size();
    blocks.reserve(std::max<size_t>(total, 1));
    auto push = [&](const std::vector<int> &blk, int &off, int &cnt) {
        off = (int)blocks.size();
        cnt = (int)blk.size() - 1;
        blocks.insert(blocks.end(), blk.begin(), blk.end());
    };
    for (size_t k = 0; k < count; ++k) {
        hprqp::batch::Inst &I = b->inst[k];
        if (I.dense) continue;
        push(per[3 * k], I.q_blk, I.q_nblk);
        push(per[3 * k + 1], I.a_blk, I.a_nblk);
        push(per[3 * k + 2], I.at_blk, I.at_nblk);
    }
    if (blocks.empty()) blocks.push_back(0);
    return blocks;
}

// Uploads a block plan and the instance table that points into it.
bool upload_plan(HPRQPHipBatch *b, const std::vector<int> &blocks)
{
    if (blocks.size() > b->blocks_cap) {
        if (b->d_blocks) (void)hipFree(b->d_blocks);
        b->d_blocks = nullptr;
        if (hipMalloc((void **)&b->d_blocks, sizeof(int) * blocks.size()) != hipSuccess) return false;
        b->blocks_cap = blocks.size();
        b->ar.blocks = b->d_blocks;
    }
    return hipMemcpy(b->d_blocks, blocks.data(), sizeof(int) * blocks.size(), hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(b->d_inst, b->inst.data(), sizeof(hprqp::batch::Inst) * b->count, hipMemcpyHostToDevice) ==
               hipSuccess;
}

}  // namespace

namespace {

// ---------------------------------------------------- device transposes
// The A' column indices and the A' -> A and Q' -> Q entry permutations of
// every CSR instance (the host counting-sort transpose: columns ascending,
// rows ascending within a column) from one stable device transpose of the
// block-diagonal matrix of all CSR instances: an instance's columns are one
// contiguous range of it, so its transposed entries land at its own entry
// offsets and only the row and entry numbering needs making local again.

// which = 0: A (rows m, columns n), 1: Q (rows n, columns n). off[k] =
// (row offset of A, column offset), among the CSR instances only.
__global__ __launch_bounds__(hprqp::kThreads) void batch_global_csr_kernel(hprqp::batch::Arrays ar, const int2 *off, int which,
                                                                          int *grp, int *gci, int total_rows,
                                                                          int total_nnz)
{
    const int k = blockIdx.x;
    if (k == 0 && threadIdx.x == 0) grp[total_rows] = total_nnz;
    const hprqp::batch::Inst I = ar.inst[k];
    if (I.dense) return;
    const int rows = which ? I.n : I.m;
    const int *rp = ar.rp + (which ? I.q_rp : I.a_rp);
    const long long nz = which ? I.q_nz : I.a_nz;
    const int *ci = ar.ci + (which ? 0 : ar.a_base) + nz;
    const int row_off = which ? off[k].y : off[k].x, col_off = off[k].y;
    for (int i = threadIdx.x; i < rows; i += hprqp::kThreads) grp[row_off + i] = (int)nz + rp[i];
    const int cnt = rp[rows];
    for (int e = threadIdx.x; e < cnt; e += hprqp::kThreads) gci[nz + e] = col_off + ci[e];
}

__global__ __launch_bounds__(hprqp::kThreads) void batch_localize_kernel(hprqp::batch::Arrays ar, const int2 *off, int which)
{
    const int k = blockIdx.x;
    const hprqp::batch::Inst I = ar.inst[k];
    if (I.dense) return;
    if (which == 0) {
        const int cnt = ar.rp[I.a_rp + I.m];
        int *tci = ar.ci + ar.at_ci_base + I.a_nz, *perm = ar.perm + I.a_nz;
        for (int e = threadIdx.x; e < cnt; e += hprqp::kThreads) {
            tci[e] -= off[k].x;
            perm[e] -= (int)I.a_nz;
        }
    } else {
        const int cnt = ar.rp[I.q_rp + I.n];
        int *perm = ar.perm + ar.qt_perm_base + I.q_nz;
        for (int e = threadIdx.x; e < cnt; e += hprqp::kThreads) perm[e] -= (int)I.q_nz;
    }
}

HPRQPStatus batch_device_transposes(HPRQPHipBatch *b)
{
    using namespace hprqp::batch;
    const int count = (int)b->count;
    long long rows_a = 0, cols = 0;
    for (int k = 0; k < count; ++k)
        if (!b->inst[k].dense) {
            rows_a += b->inst[k].m;
            cols += b->inst[k].n;
        }
    const long long nnz_a = b->ar.at_ci_base - b->ar.a_base, nnz_q = b->ar.a_base;
    hprqp::DeviceBuffer<int2> off;
    hprqp::DeviceBuffer<int> grp, gci, trp;
    HPRQP_HIP(off.alloc((size_t)count));
    HPRQP_HIP(hipMemcpyAsync(off.p, b->tr_off.data(), sizeof(int2) * count, hipMemcpyHostToDevice, b->stream));
    HPRQP_HIP(grp.alloc((size_t)std::max(rows_a, cols) + 1));
    HPRQP_HIP(gci.alloc((size_t)std::max(nnz_a, nnz_q)));
    HPRQP_HIP(trp.alloc((size_t)cols + 1));
    for (int which = 0; which < 2; ++which) {
        const long long rows = which ? cols : rows_a, nnz = which ? nnz_q : nnz_a;
        batch_global_csr_kernel<<<count, hprqp::kThreads, 0, b->stream>>>(b->ar, off.p, which, grp.p, gci.p, (int)rows,
                                                                          (int)nnz);
        HPRQP_HIP(hipGetLastError());
        int *tci = which ? nullptr : b->ar.ci + b->ar.at_ci_base;
        int *perm = which ? b->ar.perm + b->ar.qt_perm_base : b->ar.perm;
        const HPRQPStatus st = hprqp_hip_csr_transpose((int)rows, (int)cols, (int)nnz, grp.p, gci.p, nullptr, trp.p,
                                                       tci, nullptr, perm, b->stream);
        if (st != HPRQP_STATUS_OK) return st;
        batch_localize_kernel<<<count, hprqp::kThreads, 0, b->stream>>>(b->ar, off.p, which);
        HPRQP_HIP(hipGetLastError());
    }
    HPRQP_HIP(hipStreamSynchronize(b->stream));
    return HPRQP_STATUS_OK;
}

// HPRQP_BATCH_TRACE=1 reports the stages of batch creation on stderr
// (diagnostics): wall time since the previous mark.
struct StageClock {
    bool on = std::getenv("HPRQP_BATCH_TRACE") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(const char *what)
    {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "hprqp batch: %-28s %9.3f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

// Common end of batch creation: instance table upload, state and stream,
// then the setup kernels (per-instance scaling, power iteration, sigma0,
// iteration-0 residuals). Deletes the batch on failure.
HPRQPStatus batch_finish(HPRQPHipBatch *b)
{
    StageClock clock;
    using namespace hprqp::batch;
    const size_t count = b->count;
    // largest instances first, so the short ones fill the tail
    std::vector<int> order(count);
    for (size_t k = 0; k < count; ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->work[a] > b->work[c]; });
    HPRQPStatus st = HPRQP_STATUS_OK;
    if (hipMalloc((void **)&b->d_inst, sizeof(Inst) * count) != hipSuccess ||
        hipMemcpy(b->d_inst, b->inst.data(), sizeof(Inst) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc((void **)&b->d_order, sizeof(int) * count) != hipSuccess ||
        hipMemcpy(b->d_order, order.data(), sizeof(int) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc((void **)&b->d_state, sizeof(State) * count) != hipSuccess ||
        hipMalloc((void **)&b->d_active, sizeof(int)) != hipSuccess ||
        (!b->stream && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess))
        st = HPRQP_STATUS_ALLOCATION_FAILED;
    if (st != HPRQP_STATUS_OK) {
        delete b;
        return st;
    }
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            b->lone_max = 3 * cus;
    }
    Arrays &ar = b->ar;
    ar.inst = b->d_inst;
    ar.state = b->d_state;
    ar.active = b->d_active;
    ar.order = b->d_order;
    ar.wave_row = kWaveRowDefault;
    for (size_t k = 0; k < count; ++k) (b->inst[k].dense ? b->any_dense : b->any_sparse) = true;
    b->stage_x = !std::getenv("HPRQP_BATCH_NO_STAGE");
    for (size_t k = 0; k < count; ++k)
        if (!b->inst[k].dense && (b->inst[k].n > kStageMax || b->inst[k].m > kStageMax)) b->stage_x = false;
    if (b->any_dense) {
        // G = A A^T per dense instance (m x m each)
        long long gsum = 0;
        for (size_t k = 0; k < count; ++k)
            if (b->inst[k].dense) {
                b->inst[k].dg = gsum;
                gsum += (long long)b->inst[k].m * b->inst[k].m;
            }
        if (hipMalloc((void **)&b->d_gram, sizeof(double) * (size_t)std::max(gsum, 1LL)) != hipSuccess ||
            hipMemcpy(b->d_inst, b->inst.data(), sizeof(Inst) * count, hipMemcpyHostToDevice) != hipSuccess) {
            delete b;
            return HPRQP_STATUS_ALLOCATION_FAILED;
        }
        ar.gram = b->d_gram;
    }
    clock.mark("device state and order");
    if (!b->tr_off.empty()) {
        const HPRQPStatus tst = batch_device_transposes(b);
        if (tst != HPRQP_STATUS_OK) {
            delete b;
            return tst;
        }
        clock.mark("transposes (device)");
    }
    if (b->any_sparse) {
        bool fits = true;
        for (size_t k = 0; k < count; ++k)
            if (!b->inst[k].dense && (b->inst[k].n > kPowMax || b->inst[k].m > kPowMax)) fits = false;
        if (fits) {
            batch_setup_kernel<false><<<(unsigned)count, hprqp::kThreads, 0, b->stream>>>(ar, b->P, (int)count, 1);
            if (clock.on && hipStreamSynchronize(b->stream) == hipSuccess) clock.mark("setup: scaling (CSR)");
            batch_power_kernel<<<(unsigned)count, kPowThreads, 0, b->stream>>>(ar, (int)count);
            if (clock.on && hipStreamSynchronize(b->stream) == hipSuccess) clock.mark("setup: power iterations (CSR)");
            batch_setup_kernel<false><<<(unsigned)count, hprqp::kThreads, 0, b->stream>>>(ar, b->P, (int)count, 3);
        } else {
            batch_setup_kernel<false><<<(unsigned)count, hprqp::kThreads, 0, b->stream>>>(ar, b->P, (int)count, 0);
        }
    }
    if (b->any_dense) {
        batch_setup_kernel<true><<<(unsigned)count, hprqp::kThreads, 0, b->stream>>>(ar, b->P, (int)count, 1);
        if (clock.on && hipStreamSynchronize(b->stream) == hipSuccess) clock.mark("setup: scaling (dense)");
        dense_gram_kernel<<<(unsigned)count, hprqp::kThreads, 0, b->stream>>>(ar, (int)count);
        if (clock.on && hipStreamSynchronize(b->stream) == hipSuccess) clock.mark("setup: G = A A' (dense)");
        batch_setup_kernel<true><<<(unsigned)count, hprqp::kThreads, 0, b->stream>>>(ar, b->P, (int)count, 2);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) {
        delete b;
        return HPRQP_STATUS_NOT_IMPLEMENTED;
    }
    clock.mark("setup: last phase (residuals; dense: power iterations)");
    if (b->any_sparse) {
        // Wave threshold for the CSR engine: HPRQP_BATCH_WAVE_ROW, or the
        // fastest of a few candidates on a timed probe of one iteration's
        // products (the plan changes only the order of long-row sums).
        const char *force = std::getenv("HPRQP_BATCH_WAVE_ROW");
        std::vector<int> cands = {16, 32, 64, 128, 1 << 20};
        if (!b->max_row.empty()) {
            // thresholds above every row length give one and the same plan
            const int longest = *std::max_element(b->max_row.begin(), b->max_row.end());
            std::vector<int> kept;
            for (int T : cands)
                if (T <= longest) kept.push_back(T);
            kept.push_back(1 << 20);
            cands = kept;
        }
        if (force && std::atoi(force) > 0) cands = {std::atoi(force)};
        int best = kWaveRowDefault;
        float best_ms = 1e30f;
        std::vector<int> best_plan;
        std::vector<Inst> best_inst;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
        for (size_t c = 0; ok && c < cands.size(); ++c) {
            const int T = cands[c];
            std::vector<int> plan = plan_all_blocks(b, b->h_rp.get(), T);
            ok = upload_plan(b, plan);
            if (!ok || cands.size() == 1) {
                best = T;
                best_plan = std::move(plan);
                best_inst = b->inst;
                break;
            }
            ar.wave_row = T;
            float ms_min = 1e30f;
            for (int rep = 0; ok && rep < 4; ++rep) {
                ok = hipEventRecord(e0, b->stream) == hipSuccess;
                batch_probe_kernel<<<(unsigned)count, hprqp::kThreads, 0, b->stream>>>(ar, b->P, (int)count);
                ok = ok && hipEventRecord(e1, b->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
                float ms = 0.0f;
                ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
                if (rep > 0) ms_min = std::min(ms_min, ms);  // first launch warms up
            }
            if (ms_min < best_ms) {
                best_ms = ms_min;
                best = T;
                best_plan = std::move(plan);
                best_inst = b->inst;
            }
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        ar.wave_row = best;
        if (ok) b->inst = std::move(best_inst);
        if (!ok || !upload_plan(b, best_plan) || hipGetLastError() != hipSuccess) {
            delete b;
            return HPRQP_STATUS_NOT_IMPLEMENTED;
        }
        clock.mark("wave threshold probes");
    }
    return HPRQP_STATUS_OK;
}

}  // namespace

extern "C" {

HPRQPStatus hprqp_hip_batch_create(const HPRQPModel *models, size_t count, const HPRQPParams *params,
                                   HPRQPHipBatch **out)
{
    using namespace hprqp::batch;
    if (!models || !params || !out || count == 0 || count > (size_t)INT32_MAX) return HPRQP_STATUS_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return HPRQP_STATUS_NOT_IMPLEMENTED;
    if (params->device_number < 0 || params->device_number >= ndev) return HPRQP_STATUS_INVALID_ARGUMENT;
    StageClock clock;
    HPRQP_HIP(hipSetDevice(params->device_number));
    auto *b = new (std::nothrow) HPRQPHipBatch();
    if (!b) return HPRQP_STATUS_ALLOCATION_FAILED;
    clock.mark("device selection");
    b->count = count;
    b->time_limit = params->time_limit;
    b->P = Params{params->stoptol, (long long)std::min(params->max_iter, (size_t)INT64_MAX),
                  (long long)params->check_iter, params->print_frequency, params->use_scaling};
    // sizes and validation
    long long sum_n = 0, sum_m = 0;
    for (size_t k = 0; k < count; ++k) {
        const HPRQPModel &md = models[k];
        if (md.num_var_cones || md.num_row_cones || md.Q.ncol == 0 || !matrix_csr_is_valid(&md.Q) ||
            !matrix_csr_is_valid(&md.A)) {
            delete b;
            return md.num_var_cones || md.num_row_cones ? HPRQP_STATUS_NOT_IMPLEMENTED : HPRQP_STATUS_INVALID_ARGUMENT;
        }
        sum_n += (long long)md.Q.ncol;
        sum_m += (long long)md.A.nrow;
    }
    // Layout pass: per-instance offsets (dense or CSR), so that packing can
    // run on worker threads.
    const char *dense_env = std::getenv("HPRQP_BATCH_DENSE");
    const int dense_mode = dense_env ? std::atoi(dense_env) : -1;  // -1 auto, 0 never, 1 whenever it fits
    b->inst.resize(count);
    long long sum_q = 0, sum_a = 0, sum_rp = 0, sum_d = 0, vno = 0, vmo = 0;
    for (size_t k = 0; k < count; ++k) {
        const HPRQPModel &md = models[k];
        Inst &I = b->inst[k];
        const long long n = (long long)md.Q.ncol, m = (long long)md.A.nrow;
        I = Inst{};
        I.n = (int)n;
        I.m = (int)m;
        I.vn = vno;
        I.vm = vmo;
        const bool fits = n <= kDenseMax && m <= kDenseMax && m > 0;
        const double dense_bytes = 8.0 * (double)(n * n + 2 * m * n);
        const double csr_bytes = 12.0 * (double)(md.Q.nnz + 2 * md.A.nnz);
        I.dense = fits && (dense_mode == 1 || (dense_mode < 0 && dense_bytes <= csr_bytes)) ? 1 : 0;
        if (I.dense) {
            I.dq = sum_d;
            I.dat = sum_d + n * n;
            I.da = I.dat + m * n;
            sum_d += n * n + 2 * m * n;
        } else {
            I.q_nz = sum_q;
            I.a_nz = sum_a;
            I.q_rp = sum_rp;
            I.a_rp = I.q_rp + n + 1;
            I.at_rp = I.a_rp + m + 1;
            I.qt_rp = I.at_rp + n + 1;
            sum_rp += 3 * (n + 1) + m + 1;
            sum_q += (long long)md.Q.nnz;
            sum_a += (long long)md.A.nnz;
        }
        vno += n;
        vmo += m;
    }
    if (sum_q + 2 * sum_a > INT32_MAX || sum_rp > INT32_MAX) {
        delete b;
        return HPRQP_STATUS_INVALID_ARGUMENT;
    }
    const long long a_base = sum_q, at_base = sum_q + sum_a;
    // Host staging, written once by the packing threads (no zero fill: every
    // entry is written by exactly one instance's pack). The work vectors are
    // cleared on the device; only the five input vectors travel.
    const size_t n_rp = (size_t)std::max(sum_rp, 1LL), n_nz = (size_t)std::max(sum_q + 2 * sum_a, 1LL);
    const size_t n_perm = (size_t)std::max(sum_a + sum_q, 1LL), n_d = (size_t)std::max(sum_d, 1LL);
    const size_t n_host = (size_t)std::max(at_base, 1LL);  // Q and A entries; A' is formed on the device
    const size_t sm1 = (size_t)std::max(sum_m, 1LL);
    HostArray<int> ci = host_array<int>(n_host), rp = host_array<int>(n_rp);
    HostArray<double> val = host_array<double>(n_host), dval = host_array<double>(n_d);
    HostArray<double> in_n = host_array<double>((size_t)sum_n * 3), in_m = host_array<double>(sm1 * 2);
    if (!ci || !val || !dval || !in_n || !in_m || !rp) {
        delete b;
        return HPRQP_STATUS_ALLOCATION_FAILED;
    }
    // placeholders of empty arrays (one entry each) are written here
    ci[0] = rp[0] = 0;
    val[0] = dval[0] = in_m[0] = in_m[1] = 0.0;
    b->max_row.assign(count, 0);
    clock.mark("layout and staging");
    auto pack = [&](size_t k) {
        const HPRQPModel &md = models[k];
        Inst &I = b->inst[k];
        const size_t n = md.Q.ncol, m = md.A.nrow;
        {
            size_t free_count = 0;
            for (size_t i = 0; i < n; ++i)
                if (std::isinf(md.l[i]) && std::isinf(md.u[i])) ++free_count;
            I.noC = 5 * free_count > 4 * n ? 1 : 0;  // workspace.c:45-61
        }
        I.obj_constant = md.obj_constant;
        double nb = 0.0, nc = 0.0;
        for (size_t i = 0; i < m; ++i) {
            double a = std::isinf(md.AL[i]) ? 0.0 : std::fabs(md.AL[i]);
            double u = std::isinf(md.AU[i]) ? 0.0 : std::fabs(md.AU[i]);
            nb = std::max(nb, std::max(a, u));
        }
        for (size_t i = 0; i < n; ++i) nc = std::max(nc, std::fabs(md.c[i]));
        I.norm_b_org = nb;
        I.norm_c_org = nc;
        if (I.dense) {
            // Q and A row-major (zeros where the CSR has no entry); the A^T
            // copy is written on the device after scaling
            double *Qd = dval.get() + I.dq, *Ad = dval.get() + I.dat;
            std::fill(Qd, Qd + n * n, 0.0);
            std::fill(Ad, Ad + 2 * m * n, 0.0);  // A, then its transposed copy (written on the device)
            for (size_t r = 0; r < n; ++r)
                for (size_t e = md.Q.row_ptr[r]; e < md.Q.row_ptr[r + 1]; ++e) Qd[r * n + md.Q.col_ind[e]] += md.Q.values[e];
            for (size_t r = 0; r < m; ++r)
                for (size_t e = md.A.row_ptr[r]; e < md.A.row_ptr[r + 1]; ++e) Ad[r * n + md.A.col_ind[e]] += md.A.values[e];
        } else {
            // Q and A as they are; A' and Q' row pointers by column counts
            // (their entries and permutations are formed on the device)
            int *qrp = rp.get() + I.q_rp, *arp = rp.get() + I.a_rp;
            int *atrp = rp.get() + I.at_rp, *qtrp = rp.get() + I.qt_rp;
            for (size_t r = 0; r <= n; ++r) qrp[r] = (int)md.Q.row_ptr[r];
            for (size_t r = 0; r <= m; ++r) arp[r] = (int)md.A.row_ptr[r];
            std::fill(atrp, atrp + n + 1, 0);
            std::fill(qtrp, qtrp + n + 1, 0);
            int *qci = ci.get() + I.q_nz, *aci = ci.get() + a_base + I.a_nz;
            double *qv = val.get() + I.q_nz, *av = val.get() + a_base + I.a_nz;
            for (size_t e = 0; e < md.Q.nnz; ++e) {
                const int c = (int)md.Q.col_ind[e];
                qci[e] = c;
                qv[e] = md.Q.values[e];
                ++qtrp[c + 1];
            }
            for (size_t e = 0; e < md.A.nnz; ++e) {
                const int c = (int)md.A.col_ind[e];
                aci[e] = c;
                av[e] = md.A.values[e];
                ++atrp[c + 1];
            }
            int longest = 0;
            for (size_t j = 0; j < n; ++j) {
                longest = std::max({longest, atrp[j + 1], qrp[j + 1] - qrp[j]});
                atrp[j + 1] += atrp[j];
                qtrp[j + 1] += qtrp[j];
            }
            for (size_t r = 0; r < m; ++r) longest = std::max(longest, arp[r + 1] - arp[r]);
            b->max_row[k] = longest;
        }
        for (size_t i = 0; i < n; ++i) {
            in_n[I.vn + i] = md.c[i];
            in_n[(size_t)sum_n + I.vn + i] = md.l[i];
            in_n[2 * (size_t)sum_n + I.vn + i] = md.u[i];
        }
        for (size_t i = 0; i < m; ++i) {
            in_m[I.vm + i] = md.AL[i];
            in_m[sm1 + I.vm + i] = md.AU[i];
        }
    };
    {
        const size_t workers = std::max<size_t>(1, std::min<size_t>({16, count, (size_t)std::max(1u, std::thread::hardware_concurrency())}));
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        for (size_t w = 1; w < workers; ++w)
            pool.emplace_back([&] {
                for (size_t k; (k = next.fetch_add(1)) < count;) pack(k);
            });
        for (size_t k; (k = next.fetch_add(1)) < count;) pack(k);
        for (std::thread &th : pool) th.join();
    }
    for (size_t k = 0; k < count; ++k) {
        b->n.push_back(models[k].Q.ncol);
        b->m.push_back(models[k].A.nrow);
        b->work.push_back((double)(models[k].Q.nnz + 2 * models[k].A.nnz + models[k].Q.ncol + models[k].A.nrow));
    }
    clock.mark("packing (threads)");
    std::vector<int> blocks = plan_all_blocks(b, rp.get(), kWaveRowDefault);
    clock.mark("row-block plan");
    HPRQPStatus st = HPRQP_STATUS_OK;
    auto upload = [&](auto *&dst, const auto *src, size_t count) -> bool {
        using T = typename std::remove_const<typename std::remove_pointer<decltype(src)>::type>::type;
        if (hipMalloc((void **)&dst, sizeof(T) * count) != hipSuccess) return false;
        return hipMemcpy(dst, src, sizeof(T) * count, hipMemcpyHostToDevice) == hipSuccess;
    };
    auto vectors = [&](double *&dst, size_t len, size_t slots, const double *in, std::initializer_list<int> at) -> bool {
        if (hipMalloc((void **)&dst, sizeof(double) * len * slots) != hipSuccess ||
            hipMemset(dst, 0, sizeof(double) * len * slots) != hipSuccess)
            return false;
        size_t j = 0;
        for (int slot : at)
            if (hipMemcpy(dst + (size_t)slot * len, in + len * j++, sizeof(double) * len, hipMemcpyHostToDevice) !=
                hipSuccess)
                return false;
        return true;
    };
    // ci and val hold Q, A, then A' (formed on the device); perm is formed on the device
    auto upload_part = [&](auto *&dst, const auto *src, size_t count, size_t part) -> bool {
        using T = typename std::remove_const<typename std::remove_pointer<decltype(src)>::type>::type;
        if (hipMalloc((void **)&dst, sizeof(T) * count) != hipSuccess) return false;
        return hipMemcpy(dst, src, sizeof(T) * part, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!upload(b->d_rp, rp.get(), n_rp) || !upload_part(b->d_ci, ci.get(), n_nz, n_host) ||
        !upload_part(b->d_val, val.get(), n_nz, n_host) ||
        hipMalloc((void **)&b->d_perm, sizeof(int) * n_perm) != hipSuccess ||
        !upload(b->d_blocks, blocks.data(), blocks.size()) ||
        !vectors(b->d_vn, (size_t)sum_n, N_COUNT, in_n.get(), {N_c, N_l, N_u}) ||
        !vectors(b->d_vm, sm1, M_COUNT, in_m.get(), {M_AL, M_AU}) || !upload(b->d_dval, dval.get(), n_d))
        st = HPRQP_STATUS_ALLOCATION_FAILED;
    b->h_rp = std::move(rp);
    b->blocks_cap = blocks.size();
    if (st != HPRQP_STATUS_OK) {
        delete b;
        return st;
    }
    b->ar.rp = b->d_rp;
    b->ar.ci = b->d_ci;
    b->ar.val = b->d_val;
    b->ar.perm = b->d_perm;
    b->ar.blocks = b->d_blocks;
    b->ar.vn = b->d_vn;
    b->ar.vm = b->d_vm;
    b->ar.dval = b->d_dval;
    b->ar.sum_n = sum_n;
    b->ar.sum_m = std::max(sum_m, 1LL);
    b->ar.at_ci_base = at_base;
    b->ar.a_base = a_base;
    b->ar.qt_perm_base = sum_a;
    b->tr_off.resize(count);
    for (size_t k = 0, ra = 0, cn = 0; k < count; ++k) {
        b->tr_off[k] = make_int2((int)ra, (int)cn);
        if (!b->inst[k].dense) {
            ra += b->inst[k].m;
            cn += b->inst[k].n;
        }
    }
    clock.mark("upload");
    st = batch_finish(b);
    if (st == HPRQP_STATUS_OK) *out = b;
    return st;
}

// Relaunches the chunk kernel (bodies_per_launch loop bodies per instance)
// until every instance finished or the time limit passed.
HPRQPStatus hprqp_hip_batch_solve(HPRQPHipBatch *b, size_t bodies_per_launch, size_t *remaining)
{
    using namespace hprqp::batch;
    if (!b || bodies_per_launch == 0) return HPRQP_STATUS_INVALID_ARGUMENT;
    const double t0 = hprqp::wall_seconds();
    // One chunk = clear the counter, run `bodies_per_launch` loop bodies of
    // every unfinished instance (CSR and dense kernels). It is captured once
    // into a graph and replayed until no instance is left (HPRQP_BATCH_GRAPH=0
    // launches it eagerly instead).
    const char *graph_env = std::getenv("HPRQP_BATCH_GRAPH");
    const bool use_graph = !(graph_env && std::atoi(graph_env) == 0);
    // The tail: with at most lone_max unfinished instances the CSR instances
    // continue in batch_solve_lone_kernel. HPRQP_BATCH_LONE=0 keeps the
    // shared kernel to the end, =1 runs the lone kernel from the start;
    // HPRQP_BATCH_LONE_MAX overrides the threshold.
    const char *lone_env = std::getenv("HPRQP_BATCH_LONE");
    const int lone_mode = lone_env ? std::atoi(lone_env) : -1;
    const char *lone_max_env = std::getenv("HPRQP_BATCH_LONE_MAX");
    const int lone_max = lone_max_env && std::atoi(lone_max_env) > 0 ? std::atoi(lone_max_env) : b->lone_max;
    auto enqueue_chunk = [&](bool lone) {
        batch_zero_counter_kernel<<<1, 1, 0, b->stream>>>(b->d_active);
        if (b->any_sparse && lone && b->stage_x)
            batch_solve_lone_kernel<true><<<(unsigned)b->count, kLoneThreads, 0, b->stream>>>(
                b->ar, b->P, (int)b->count, (int)bodies_per_launch);
        else if (b->any_sparse && lone)
            batch_solve_lone_kernel<false><<<(unsigned)b->count, kLoneThreads, 0, b->stream>>>(
                b->ar, b->P, (int)b->count, (int)bodies_per_launch);
        else if (b->any_sparse && b->stage_x)
            batch_solve_kernel<false, true><<<(unsigned)b->count, hprqp::kThreads, 0, b->stream>>>(
                b->ar, b->P, (int)b->count, (int)bodies_per_launch);
        else if (b->any_sparse)
            batch_solve_kernel<false, false><<<(unsigned)b->count, hprqp::kThreads, 0, b->stream>>>(
                b->ar, b->P, (int)b->count, (int)bodies_per_launch);
        if (b->any_dense)
            batch_solve_kernel<true, false><<<(unsigned)b->count, hprqp::kThreads, 0, b->stream>>>(
                b->ar, b->P, (int)b->count, (int)bodies_per_launch);
    };
    auto capture = [&](bool lone, hipGraphExec_t *exec, size_t *bodies) -> HPRQPStatus {
        if (*exec && *bodies == bodies_per_launch) return HPRQP_STATUS_OK;
        if (*exec) HPRQP_HIP(hipGraphExecDestroy(*exec));
        *exec = nullptr;
        hipGraph_t g = nullptr;
        HPRQP_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        enqueue_chunk(lone);
        HPRQP_HIP(hipStreamEndCapture(b->stream, &g));
        const hipError_t ie = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HPRQP_HIP(ie);
        *bodies = bodies_per_launch;
        return HPRQP_STATUS_OK;
    };
    int active = (int)b->count;
    while (active > 0) {
        const bool lone = b->any_sparse && lone_mode != 0 && (lone_mode == 1 || active <= lone_max);
        if (use_graph) {
            const HPRQPStatus cs = lone ? capture(true, &b->graph_lone, &b->graph_lone_bodies)
                                        : capture(false, &b->graph, &b->graph_bodies);
            if (cs != HPRQP_STATUS_OK) return cs;
            HPRQP_HIP(hipGraphLaunch(lone ? b->graph_lone : b->graph, b->stream));
        } else {
            enqueue_chunk(lone);
        }
        HPRQP_HIP(hipGetLastError());
        HPRQP_HIP(hipMemcpyAsync(&active, b->d_active, sizeof(int), hipMemcpyDeviceToHost, b->stream));
        HPRQP_HIP(hipStreamSynchronize(b->stream));
        ++b->launches;
        if (hprqp::wall_seconds() - t0 > b->time_limit) break;
    }
    b->solve_seconds = hprqp::wall_seconds() - t0;
    b->state.resize(b->count);
    HPRQP_HIP(hipMemcpy(b->state.data(), b->d_state, sizeof(State) * b->count, hipMemcpyDeviceToHost));
    if (remaining) *remaining = (size_t)active;
    return HPRQP_STATUS_OK;
}

double hprqp_hip_batch_seconds(const HPRQPHipBatch *b) { return b ? b->solve_seconds : 0.0; }

int hprqp_hip_batch_wave_row(const HPRQPHipBatch *b) { return b ? b->ar.wave_row : 0; }

// Instance k's results in original units (status TIME_LIMIT if it did not
// finish before the limit).
HPRQPStatus hprqp_hip_batch_result(const HPRQPHipBatch *b, size_t k, HPRQPResults *res)
{
    using namespace hprqp::batch;
    if (!b || !res || k >= b->count || b->state.size() != b->count) return HPRQP_STATUS_INVALID_ARGUMENT;
    const State &S = b->state[k];
    const Inst &I = b->inst[k];
    const size_t n = b->n[k], m = b->m[k];
    results_free(res);
    res->x = static_cast<double *>(std::calloc(n ? n : 1, sizeof(double)));
    res->w = static_cast<double *>(std::calloc(n ? n : 1, sizeof(double)));
    res->z = static_cast<double *>(std::calloc(n ? n : 1, sizeof(double)));
    res->y = static_cast<double *>(std::calloc(m ? m : 1, sizeof(double)));
    std::vector<double> xb(n), wb(n), zb(n), cn(n), yb(m), rn(m);
    auto get = [&](double *dst, const double *base, long long vec, long long len_total, long long off, size_t len) {
        if (len) (void)hipMemcpy(dst, base + vec * len_total + off, sizeof(double) * len, hipMemcpyDeviceToHost);
    };
    get(xb.data(), b->d_vn, N_x_bar, b->ar.sum_n, I.vn, n);
    get(wb.data(), b->d_vn, N_w_bar, b->ar.sum_n, I.vn, n);
    get(zb.data(), b->d_vn, N_z_bar, b->ar.sum_n, I.vn, n);
    get(cn.data(), b->d_vn, N_col, b->ar.sum_n, I.vn, n);
    get(yb.data(), b->d_vm, M_y_bar, b->ar.sum_m, I.vm, m);
    get(rn.data(), b->d_vm, M_row, b->ar.sum_m, I.vm, m);
    for (size_t i = 0; i < n; ++i) {
        res->x[i] = xb[i] / cn[i];
        res->w[i] = wb[i] / cn[i];
        res->z[i] = zb[i] * cn[i];
    }
    for (size_t i = 0; i < m; ++i) res->y[i] = yb[i] / rn[i];
    res->iter = (size_t)S.iter;
    res->iter_4 = S.first4 ? (size_t)S.iter : (size_t)S.iter4;
    res->iter_6 = S.first6 ? (size_t)S.iter : (size_t)S.iter6;
    res->time = b->solve_seconds;
    res->primal_obj = S.pobj;
    res->residuals = S.kkt;
    res->gap = S.gap;
    res->status = S.finished ? S.status : HPRQP_STATUS_TIME_LIMIT;
    return HPRQP_STATUS_OK;
}

}  // extern "C"

namespace {

// Shared by the host-array and device-array dense entry points. With
// on_device the arrays live on the batch's device and are read on `caller`
// (the stream that produced them), which is synchronised before returning;
// otherwise they are host arrays.
HPRQPStatus create_dense_impl(size_t count, size_t n, size_t m, const double *Q, const double *c, const double *A,
                              const double *AL, const double *AU, const double *l, const double *u,
                              const HPRQPParams *params, bool on_device, hipStream_t caller, HPRQPHipBatch **out)
{
    using namespace hprqp::batch;
    if (!out || !params || !Q || !c || !l || !u || (m && (!A || !AL || !AU)) || count == 0 || n == 0 || m == 0 ||
        n > (size_t)kDenseMax || m > (size_t)kDenseMax || count > (size_t)INT32_MAX / (n > m ? n : m))
        return HPRQP_STATUS_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return HPRQP_STATUS_NOT_IMPLEMENTED;
    if (params->device_number < 0 || params->device_number >= ndev) return HPRQP_STATUS_INVALID_ARGUMENT;
    HPRQP_HIP(hipSetDevice(params->device_number));
    const long long N = (long long)n, M = (long long)m, C = (long long)count;
    const long long sum_n = C * N, sum_m = C * M;
    std::vector<DenseStats> stats(count);
    if (on_device) {
        // the inputs must live on this device
        for (const double *p : {Q, c, A, AL, AU, l, u}) {
            hipPointerAttribute_t at{};
            if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type != hipMemoryTypeDevice ||
                at.device != params->device_number) {
                (void)hipGetLastError();
                return HPRQP_STATUS_INVALID_ARGUMENT;
            }
        }
        DenseStats *d_stats = nullptr;
        HPRQP_HIP(hipMalloc((void **)&d_stats, sizeof(DenseStats) * count));
        dense_inspect_kernel<<<(unsigned)count, hprqp::kThreads, 0, caller>>>(Q, c, A, AL, AU, l, u, (int)n, (int)m,
                                                                          d_stats);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(stats.data(), d_stats, sizeof(DenseStats) * count, hipMemcpyDeviceToHost,
                                                caller);
        if (e == hipSuccess) e = hipStreamSynchronize(caller);
        (void)hipFree(d_stats);
        HPRQP_HIP(e);
    } else {
        // validation (build.c:339-400 checks, plus: no row of A to delete)
        const size_t workers = std::max<size_t>(1, std::min<size_t>({16, count, (size_t)std::max(1u, std::thread::hardware_concurrency())}));
        std::atomic<size_t> next{0};
        auto inspect = [&] {
            for (size_t k; (k = next.fetch_add(1)) < count;) {
                DenseStats &st = stats[k];
                st = DenseStats{0, 0, 0.0, 0.0};
                const double *Qk = Q + k * n * n, *Ak = A + k * m * n;
                for (size_t e = 0; e < n * n; ++e)
                    if (!std::isfinite(Qk[e])) st.bad = 1;
                for (size_t i = 0; i < m; ++i) {
                    bool nz = false;
                    for (size_t j = 0; j < n; ++j) {
                        const double v = Ak[i * n + j];
                        if (!std::isfinite(v)) st.bad = 1;
                        nz |= v != 0.0;
                    }
                    const double al = AL[k * m + i], au = AU[k * m + i];
                    if (!nz || std::isnan(al) || std::isnan(au) || al > au ||
                        (std::isinf(al) && al < 0 && std::isinf(au) && au > 0))
                        st.bad = 1;
                    st.norm_b = std::max(st.norm_b, std::max(std::isinf(al) ? 0.0 : std::fabs(al),
                                                             std::isinf(au) ? 0.0 : std::fabs(au)));
                }
                for (size_t j = 0; j < n; ++j) {
                    const double cj = c[k * n + j], lj = l[k * n + j], uj = u[k * n + j];
                    if (!std::isfinite(cj) || std::isnan(lj) || std::isnan(uj) || lj > uj) st.bad = 1;
                    if (std::isinf(lj) && std::isinf(uj)) ++st.free_count;
                    st.norm_c = std::max(st.norm_c, std::fabs(cj));
                }
                if (st.bad) return;  // the instance rejects the batch; others may stop early too
            }
        };
        std::vector<std::thread> pool;
        for (size_t w = 1; w < workers; ++w) pool.emplace_back(inspect);
        inspect();
        for (std::thread &th : pool) th.join();
    }
    for (const DenseStats &st : stats)
        if (st.bad) return HPRQP_STATUS_INVALID_ARGUMENT;
    auto *b = new (std::nothrow) HPRQPHipBatch();
    if (!b) return HPRQP_STATUS_ALLOCATION_FAILED;
    b->count = count;
    b->time_limit = params->time_limit;
    b->P = Params{params->stoptol, (long long)std::min(params->max_iter, (size_t)INT64_MAX),
                  (long long)params->check_iter, params->print_frequency, params->use_scaling};
    b->inst.resize(count);
    for (size_t k = 0; k < count; ++k) {
        Inst &I = b->inst[k];
        I = Inst{};
        I.n = (int)n;
        I.m = (int)m;
        I.vn = (long long)k * N;
        I.vm = (long long)k * M;
        I.dense = 1;
        I.dq = (long long)k * N * N;
        I.dat = C * N * N + (long long)k * M * N;
        I.da = C * N * N + C * M * N + (long long)k * M * N;
        I.noC = 5 * (size_t)stats[k].free_count > 4 * n ? 1 : 0;  // workspace.c:45-61
        I.norm_b_org = stats[k].norm_b;
        I.norm_c_org = stats[k].norm_c;
        b->n.push_back(n);
        b->m.push_back(m);
        b->work.push_back((double)(n * n + 2 * m * n));
    }
    b->blocks_cap = 1;
    const size_t dsz = (size_t)(C * N * N + 2 * C * M * N);
    auto fail = [&]() {
        delete b;
        return HPRQP_STATUS_ALLOCATION_FAILED;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail();
    if (hipMalloc((void **)&b->d_dval, sizeof(double) * dsz) != hipSuccess ||
        hipMalloc((void **)&b->d_vn, sizeof(double) * (size_t)sum_n * N_COUNT) != hipSuccess ||
        hipMalloc((void **)&b->d_vm, sizeof(double) * (size_t)sum_m * M_COUNT) != hipSuccess ||
        hipMalloc((void **)&b->d_rp, sizeof(int)) != hipSuccess || hipMalloc((void **)&b->d_ci, sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&b->d_val, sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&b->d_perm, sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&b->d_blocks, sizeof(int)) != hipSuccess)
        return fail();
    hipStream_t s = b->stream;
    // device inputs are copied on the caller's stream (ordered after their
    // producer) and that stream is synchronised before the batch stream uses them
    hipStream_t cs = on_device ? caller : s;
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    bool ok = hipMemsetAsync(b->d_vn, 0, sizeof(double) * (size_t)sum_n * N_COUNT, s) == hipSuccess &&
              hipMemsetAsync(b->d_vm, 0, sizeof(double) * (size_t)sum_m * M_COUNT, s) == hipSuccess &&
              (!on_device || hipStreamSynchronize(s) == hipSuccess) &&
              hipMemcpyAsync(b->d_dval, Q, sizeof(double) * (size_t)(C * N * N), kind, cs) == hipSuccess &&
              hipMemcpyAsync(b->d_dval + C * N * N, A, sizeof(double) * (size_t)(C * M * N), kind, cs) == hipSuccess &&
              hipMemcpyAsync(b->d_vn + N_c * sum_n, c, sizeof(double) * (size_t)sum_n, kind, cs) == hipSuccess &&
              hipMemcpyAsync(b->d_vn + N_l * sum_n, l, sizeof(double) * (size_t)sum_n, kind, cs) == hipSuccess &&
              hipMemcpyAsync(b->d_vn + N_u * sum_n, u, sizeof(double) * (size_t)sum_n, kind, cs) == hipSuccess &&
              hipMemcpyAsync(b->d_vm + M_AL * sum_m, AL, sizeof(double) * (size_t)sum_m, kind, cs) == hipSuccess &&
              hipMemcpyAsync(b->d_vm + M_AU * sum_m, AU, sizeof(double) * (size_t)sum_m, kind, cs) == hipSuccess &&
              (!on_device || hipStreamSynchronize(cs) == hipSuccess);
    if (!ok) return fail();
    dense_symmetrize_kernel<<<(unsigned)(C * N), hprqp::kThreads, 0, s>>>(b->d_dval, (int)n, (int)count);
    if (hipGetLastError() != hipSuccess) return fail();
    b->ar.rp = b->d_rp;
    b->ar.ci = b->d_ci;
    b->ar.val = b->d_val;
    b->ar.perm = b->d_perm;
    b->ar.blocks = b->d_blocks;
    b->ar.vn = b->d_vn;
    b->ar.vm = b->d_vm;
    b->ar.dval = b->d_dval;
    b->ar.sum_n = sum_n;
    b->ar.sum_m = sum_m;
    HPRQPStatus st = batch_finish(b);
    if (st == HPRQP_STATUS_OK) *out = b;
    return st;
}

}  // namespace

extern "C" {

// Batch of `count` QPs of one shape given as dense row-major arrays: Q
// (count x n x n), A (count x m x n), c, l, u (count x n), AL, AU (count x m).
// Equivalent to build_from_QAbc on each instance (Q symmetrised on the
// device) and hprqp_hip_batch_create, without building host models; every
// row of A needs a nonzero entry and a finite bound.
HPRQPStatus hprqp_hip_batch_create_dense(size_t count, size_t n, size_t m, const double *Q, const double *c,
                                         const double *A, const double *AL, const double *AU, const double *l,
                                         const double *u, const HPRQPParams *params, HPRQPHipBatch **out)
{
    return create_dense_impl(count, n, m, Q, c, A, AL, AU, l, u, params, false, nullptr, out);
}

// The same from arrays already in device memory of params->device_number
// (e.g. tensors of a training step): validated and copied on the device,
// read on `stream` (a hipStream_t; null = the null stream), which is
// synchronised before the call returns.
HPRQPStatus hprqp_hip_batch_create_dense_device(size_t count, size_t n, size_t m, const double *Q, const double *c,
                                                const double *A, const double *AL, const double *AU, const double *l,
                                                const double *u, const HPRQPParams *params, void *stream,
                                                HPRQPHipBatch **out)
{
    return create_dense_impl(count, n, m, Q, c, A, AL, AU, l, u, params, true, (hipStream_t)stream, out);
}

// Every instance's solution in original units, instances concatenated:
// x, z (sum of n), y (sum of m); status, iterations, primal objective and
// KKT residual per instance. Any output pointer may be null.
HPRQPStatus hprqp_hip_batch_solution(const HPRQPHipBatch *b, double *x, double *y, double *z, int *status,
                                     size_t *iters, double *objective, double *kkt)
{
    using namespace hprqp::batch;
    if (!b || b->state.size() != b->count) return HPRQP_STATUS_INVALID_ARGUMENT;
    const size_t sn = (size_t)b->ar.sum_n, sm = (size_t)b->ar.sum_m;
    std::vector<double> col, row;
    if (x || z) {
        col.resize(sn);
        HPRQP_HIP(hipMemcpy(col.data(), b->d_vn + (size_t)N_col * sn, sizeof(double) * sn, hipMemcpyDeviceToHost));
    }
    if (y) {
        row.resize(sm);
        HPRQP_HIP(hipMemcpy(row.data(), b->d_vm + (size_t)M_row * sm, sizeof(double) * sm, hipMemcpyDeviceToHost));
    }
    size_t real_m = 0;
    for (size_t k = 0; k < b->count; ++k) real_m += b->m[k];
    if (x) {
        HPRQP_HIP(hipMemcpy(x, b->d_vn + (size_t)N_x_bar * sn, sizeof(double) * sn, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < sn; ++i) x[i] /= col[i];
    }
    if (z) {
        HPRQP_HIP(hipMemcpy(z, b->d_vn + (size_t)N_z_bar * sn, sizeof(double) * sn, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < sn; ++i) z[i] *= col[i];
    }
    if (y && real_m) {
        HPRQP_HIP(hipMemcpy(y, b->d_vm + (size_t)M_y_bar * sm, sizeof(double) * real_m, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < real_m; ++i) y[i] /= row[i];
    }
    for (size_t k = 0; k < b->count; ++k) {
        const State &S = b->state[k];
        if (status) status[k] = S.finished ? S.status : HPRQP_STATUS_TIME_LIMIT;
        if (iters) iters[k] = (size_t)S.iter;
        if (objective) objective[k] = S.pobj;
        if (kkt) kkt[k] = S.kkt;
    }
    return HPRQP_STATUS_OK;
}

}  // extern "C"

namespace {

// Row stride (doubles, a multiple of 32) of the dense adjoint's per-wavefront
// partial row sums, and the dynamic LDS bytes they take.
int adj_dense_mld(const HPRQPHipBatch *b)
{
    size_t mx = 1;
    for (size_t k = 0; k < b->count; ++k)
        if (b->inst[k].dense) mx = std::max(mx, b->m[k]);
    return (int)((mx + 31) / 32 * 32);
}

size_t adj_dense_lds(const HPRQPHipBatch *b) { return sizeof(double) * (hprqp::kThreads / hprqp::kWave) * adj_dense_mld(b); }

// The batch-owned outputs of the adjoint besides the caller's vectors: step
// counts and final relative residuals, one per instance.
hipError_t ensure_adjoint_outputs(HPRQPHipBatch *b)
{
    if (!b->d_steps) {
        const hipError_t e = hipMalloc((void **)&b->d_steps, sizeof(int) * b->count);
        if (e != hipSuccess) return e;
    }
    if (!b->d_adj_resid) return hipMalloc((void **)&b->d_adj_resid, sizeof(double) * b->count);
    return hipSuccess;
}

// Launches the adjoint kernels (one workgroup per instance) on s.
void launch_adjoint(HPRQPHipBatch *b, hipStream_t s, const double *g, double *u, double *v_rows, double *v_bounds,
                    double tol, int max_steps)
{
    using namespace hprqp::batch;
    const unsigned grid = (unsigned)b->count;
    if (b->any_sparse)
        batch_adjoint_kernel<false><<<grid, hprqp::kThreads, 0, s>>>(b->ar, b->P, (int)b->count, g, u, v_rows,
                                                                     v_bounds, tol, max_steps, b->d_steps,
                                                                     b->d_adj_resid, 0);
    if (b->any_dense)
        batch_adjoint_kernel<true><<<grid, hprqp::kThreads, adj_dense_lds(b), s>>>(
            b->ar, b->P, (int)b->count, g, u, v_rows, v_bounds, tol, max_steps, b->d_steps, b->d_adj_resid,
            adj_dense_mld(b));
}

// Marks the end of work queued on a caller's stream that reads the batch, so
// that destroying the batch waits for it.
hipError_t mark_caller_use(const HPRQPHipBatch *b, hipStream_t s)
{
    if (!b->caller_done) {
        const hipError_t e = hipEventCreateWithFlags(&b->caller_done, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    return hipEventRecord(b->caller_done, s);
}

}  // namespace

extern "C" {

// Every instance's x, z (sum of n) and y (sum of m) in original units into
// device buffers on the batch's device (any may be null), written by a
// kernel on `stream` (a hipStream_t; null = the null stream) without a host
// synchronisation. Same values as hprqp_hip_batch_solution.
HPRQPStatus hprqp_hip_batch_solution_device(const HPRQPHipBatch *b, double *x, double *y, double *z, void *stream)
{
    using namespace hprqp::batch;
    if (!b || b->state.size() != b->count) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (!x && !y && !z) return HPRQP_STATUS_OK;
    long long real_m = 0;
    for (size_t k = 0; k < b->count; ++k) real_m += (long long)b->m[k];
    const long long len = std::max(b->ar.sum_n, real_m);
    batch_unscale_kernel<<<(unsigned)((len + hprqp::kThreads - 1) / hprqp::kThreads), hprqp::kThreads, 0,
                           (hipStream_t)stream>>>(b->d_vn, b->d_vm, b->ar.sum_n, b->ar.sum_m, real_m, x, y, z);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(mark_caller_use(b, (hipStream_t)stream));
    return HPRQP_STATUS_OK;
}

void hprqp_hip_batch_destroy(HPRQPHipBatch *b) { delete b; }

}  // extern "C"

extern "C" {

// Adjoint of every instance of a solved batch: for upstream gradients g
// (dL/dx in original units, instances concatenated) returns u (Sum n), the
// row multipliers v_rows (Sum m) and the bound multipliers v_bounds (Sum n) of
// the active-set KKT system K [u; v] = [g; 0]; steps[k] = MINRES steps used.
// Host arrays in and out; the scratch copies are freed on every return.
HPRQPStatus hprqp_hip_batch_adjoint(HPRQPHipBatch *b, const double *dl_dx, double tol, int max_steps, double *u,
                                    double *v_rows, double *v_bounds, int *steps)
{
    using namespace hprqp::batch;
    if (!b || !dl_dx || !u || !v_rows || !v_bounds || max_steps <= 0) return HPRQP_STATUS_INVALID_ARGUMENT;
    const size_t sn = (size_t)b->ar.sum_n, sm = (size_t)b->ar.sum_m;
    HPRQP_HIP(ensure_adjoint_outputs(b));
    hprqp::DeviceBuffer<double> d_g, d_u, d_vr, d_vb;
    HPRQP_HIP(d_g.alloc(sn));
    HPRQP_HIP(d_u.alloc(sn));
    HPRQP_HIP(d_vb.alloc(sn));
    HPRQP_HIP(d_vr.alloc(sm));
    HPRQP_HIP(hipMemcpy(d_g.p, dl_dx, sizeof(double) * sn, hipMemcpyHostToDevice));
    HPRQP_HIP(hipMemsetAsync(d_vr.p, 0, sizeof(double) * std::max<size_t>(sm, 1), b->stream));
    launch_adjoint(b, b->stream, d_g.p, d_u.p, d_vr.p, d_vb.p, tol, max_steps);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(hipStreamSynchronize(b->stream));
    HPRQP_HIP(hipMemcpy(u, d_u.p, sizeof(double) * sn, hipMemcpyDeviceToHost));
    HPRQP_HIP(hipMemcpy(v_bounds, d_vb.p, sizeof(double) * sn, hipMemcpyDeviceToHost));
    size_t real_m = 0;
    for (size_t k = 0; k < b->count; ++k) real_m += b->m[k];
    HPRQP_HIP(hipMemcpy(v_rows, d_vr.p, sizeof(double) * real_m, hipMemcpyDeviceToHost));
    if (steps) HPRQP_HIP(hipMemcpy(steps, b->d_steps, sizeof(int) * b->count, hipMemcpyDeviceToHost));
    return HPRQP_STATUS_OK;
}

// The adjoint with device buffers: dl_dx, u, v_bounds (sum of n) and v_rows
// (sum of m) on the batch's device, the kernels launched on `stream` (a
// hipStream_t; null = the null stream). steps (host, count entries) may be
// null; when given, the call waits for the kernels to read it back.
HPRQPStatus hprqp_hip_batch_adjoint_device(HPRQPHipBatch *b, const double *dl_dx, double tol, int max_steps, double *u,
                                           double *v_rows, double *v_bounds, int *steps, void *stream)
{
    using namespace hprqp::batch;
    if (!b || !dl_dx || !u || !v_bounds || max_steps <= 0 || b->state.size() != b->count)
        return HPRQP_STATUS_INVALID_ARGUMENT;
    size_t real_m = 0;
    for (size_t k = 0; k < b->count; ++k) real_m += b->m[k];
    if (real_m && !v_rows) return HPRQP_STATUS_INVALID_ARGUMENT;
    HPRQP_HIP(ensure_adjoint_outputs(b));
    hipStream_t s = (hipStream_t)stream;
    launch_adjoint(b, s, dl_dx, u, v_rows, v_bounds, tol, max_steps);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(mark_caller_use(b, s));
    if (steps) {
        HPRQP_HIP(hipMemcpyAsync(steps, b->d_steps, sizeof(int) * b->count, hipMemcpyDeviceToHost, s));
        HPRQP_HIP(hipStreamSynchronize(s));
    }
    return HPRQP_STATUS_OK;
}

// Per-instance final relative residual ||K [u; v] - [g; 0]|| / ||[g; 0]|| of
// the batch's last adjoint (0 where g vanished on the free variables), read
// on `stream` (the stream the device-buffer adjoint ran on; any stream after
// the host-array adjoint) and waited for.
HPRQPStatus hprqp_hip_batch_adjoint_residuals(const HPRQPHipBatch *b, double *residuals, void *stream)
{
    if (!b || !residuals) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (!b->d_adj_resid) return HPRQP_STATUS_INVALID_ARGUMENT;  // no adjoint has run
    hipStream_t s = (hipStream_t)stream;
    HPRQP_HIP(hipMemcpyAsync(residuals, b->d_adj_resid, sizeof(double) * b->count, hipMemcpyDeviceToHost, s));
    HPRQP_HIP(hipStreamSynchronize(s));
    return HPRQP_STATUS_OK;
}

}  // extern "C"
