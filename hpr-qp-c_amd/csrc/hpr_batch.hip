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
#include <cmath>
#include <cstdio>
#include <cstring>
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
// Rows of at least kWaveRow entries are grouped (up to 8 per block) and each
// is reduced by one wavefront (lanes stride the row, FMA, shuffle tree); the
// other rows go through the LDS stream mode with sequential row sums.
constexpr int kWaveRow = 64;

template <class Epi>
__device__ void wg_spmv(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ val,
                        const double *__restrict__ x, const int *__restrict__ blk, int nblk, const Epi &epi,
                        double *sums, double *maxs, double *prod)
{
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1), wid = t / kWave;
    for (int b = 0; b < nblk; ++b) {
        const int r0 = blk[b], r1 = blk[b + 1];
        const int k0 = rp[r0], k1 = rp[r1];
        if (k1 - k0 >= kWaveRow * (r1 - r0)) {
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
            for (int k = t; k < cnt; k += kThreads) prod[k] = val[k0 + k] * x[ci[k0 + k]];
            __syncthreads();
            const int r = r0 + t;
            if (r < r1) {
                double acc = 0.0;
                const int e = rp[r + 1] - k0;
                for (int k = rp[r] - k0; k < e; ++k) acc += prod[k];
                epi.row(r, acc, sums, maxs);
            }
        }
        __syncthreads();
    }
}

template <class F>
__device__ void wg_map(int len, const F &f, double *sums, double *maxs)
{
    for (int i = threadIdx.x; i < len; i += kThreads) f.row(i, sums, maxs);
    __syncthreads();
}

// Reduce per-thread terms; every thread then reads out[0 .. NS+NM).
template <int NS, int NM>
__device__ void wg_reduce(double *sums, double *maxs, double *out)
{
    block_reduce_store<NS, NM>(sums, maxs, out);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < (NS > 0 ? NS : 1); ++i) sums[i] = 0.0;
#pragma unroll
    for (int i = 0; i < (NM > 0 ? NM : 1); ++i) maxs[i] = 0.0;
}

__device__ __forceinline__ double dclamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Matrix view of one instance.
struct Mat {
    const int *rp, *ci;
    double *val;
    const int *blk;
    int nblk;
};

// One workgroup's view of its instance.
struct WG {
    const Inst *I;
    State *S;  // in LDS
    Mat Q, A, AT;
    const int *at_perm, *qt_rp, *qt_perm;
    double *a_val;
    double *vn[N_COUNT];
    double *vm[M_COUNT];
    double *prod;   // LDS staging
    double *red;    // LDS reduction results
    Params P;

    __device__ double *N(int v) const { return vn[v]; }
    __device__ double *M(int v) const { return vm[v]; }
};

// ---------------------------------------------------------- setup phases

// scale.c:231-318 on one instance (column j of A through A^T's permutation,
// column j of Q through Q^T's, so sums run in row order like the host).
__device__ void wg_scale(WG &g)
{
    const int n = g.I->n, m = g.I->m, t = threadIdx.x;
    for (int j = t; j < n; j += kThreads) g.N(N_col)[j] = 1.0;
    for (int i = t; i < m; i += kThreads) g.M(M_row)[i] = 1.0;
    __syncthreads();
    if (!g.P.use_scaling) return;
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
        wg_map(len, MapScaleDiv{{}, q, z, qn}, sums, maxs);
        if (a_matrix) {
            wg_spmv(g.AT.rp, g.AT.ci, g.AT.val, q, g.AT.blk, g.AT.nblk, EpiStore{{}, tmp}, sums, maxs, g.prod);
            wg_spmv(g.A.rp, g.A.ci, g.A.val, tmp, g.A.blk, g.A.nblk, EpiPowerZ{nullptr, z, q}, sums, maxs, g.prod);
        } else {
            wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, q, g.Q.blk, g.Q.nblk, EpiPowerZ{nullptr, z, q}, sums, maxs, g.prod);
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

// workspace.c:191-248
__device__ void wg_setup(WG &g, double *sums, double *maxs)
{
    const int n = g.I->n, m = g.I->m, t = threadIdx.x;
    wg_scale(g);
    for (long long k = t; k < (long long)g.AT.rp[n]; k += kThreads) g.AT.val[k] = g.a_val[g.at_perm[k]];
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
    wg_map(n, DiagInfo{nullptr, g.N(N_diag), g.Q.rp, g.Q.ci, g.Q.val}, sums, maxs);
    wg_reduce<0, 1>(sums, maxs, g.red);
    const int diag = n == 0 || g.red[0] <= 1e-15;
    __syncthreads();
    if (diag)
        wg_map(n, MapDiagFactors{{}, g.N(N_fact1), g.N(N_fact2), g.N(N_fact), g.N(N_fact_M), g.N(N_diag), sigma},
               sums, maxs);
    double lam_a = wg_power(g, true, sums, maxs) * 1.05;
    double lam_q = 0.0;
    if (diag) {
        for (int i = t; i < n; i += kThreads) maxs[0] = fmax(maxs[0], g.N(N_diag)[i]);
        wg_reduce<0, 1>(sums, maxs, g.red);
        lam_q = g.red[0];
        __syncthreads();
    } else {
        lam_q = wg_power(g, false, sums, maxs) * 1.05;
    }
    if (t == 0) {
        g.S->sigma = sigma;
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
__device__ void wg_residuals(WG &g, long long iter, double *sums, double *maxs)
{
    const int n = g.I->n, m = g.I->m;
    State &S = *g.S;
    wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, g.N(N_x_bar), g.Q.blk, g.Q.nblk,
            EpiResQ{nullptr, g.N(N_Qx), g.N(N_x_bar), g.N(N_c), g.N(N_z_bar)}, sums, maxs, g.prod);
    wg_reduce<3, 0>(sums, maxs, g.red);
    const double cx = g.red[0], xqx = g.red[1], zx = g.red[2];
    __syncthreads();
    wg_spmv(g.AT.rp, g.AT.ci, g.AT.val, g.M(M_y_bar), g.AT.blk, g.AT.nblk,
            EpiResAT{nullptr, g.N(N_ATdy), g.N(N_Qx), g.N(N_c), g.N(N_z_bar), g.N(N_col), 1.0}, sums, maxs, g.prod);
    wg_reduce<0, 3>(sums, maxs, g.red);
    const double rd = g.red[0], qxm = g.red[1], atym = g.red[2];
    __syncthreads();
    double ys = 0.0, rp = 0.0, axm = 0.0;
    if (m > 0) {
        wg_spmv(g.A.rp, g.A.ci, g.A.val, g.N(N_x_bar), g.A.blk, g.A.nblk,
                EpiResA{nullptr, g.M(M_pz), g.M(M_AL), g.M(M_AU), g.M(M_row), g.M(M_y_bar), g.M(M_s), 1.0}, sums,
                maxs, g.prod);
        wg_reduce<1, 2>(sums, maxs, g.red);
        ys = g.red[0];
        rp = g.red[1];
        axm = g.red[2];
        __syncthreads();
    }
    double viol = 0.0;
    if (iter == 0) {
        wg_map(n, MapBoundViolation{nullptr, g.N(N_dx), g.N(N_l), g.N(N_u), g.N(N_x_bar), g.N(N_col), 1.0}, sums,
               maxs);
        wg_reduce<0, 1>(sums, maxs, g.red);
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
__device__ double wg_m_norm(WG &g, double *sums, double *maxs)
{
    const int m = g.I->m;
    const State &S = *g.S;
    const int diag = S.Q_is_diag;
    wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, g.N(N_dw), g.Q.blk, g.Q.nblk, EpiMnQdw{nullptr, g.N(N_dQw), g.N(N_dx), g.N(N_dw)},
            sums, maxs, g.prod);
    wg_reduce<4, 0>(sums, maxs, g.red);
    const double dx_sq = g.red[0], dqw_dx = g.red[1], dw_dqw = g.red[2], dqw_sq = g.red[3];
    __syncthreads();
    const double sg = S.sigma;
    double m1 = 0.0, m3 = 0.0, m2 = (1.0 / sg) * dx_sq;
    m2 -= 2.0 * dqw_dx;
    if (m > 0) {
        wg_spmv(g.AT.rp, g.AT.ci, g.AT.val, g.M(M_dy), g.AT.blk, g.AT.nblk, EpiMnATdy{nullptr, g.N(N_ATdy), g.N(N_dx)},
                sums, maxs, g.prod);
        wg_reduce<1, 0>(sums, maxs, g.red);
        const double atdy_dx = g.red[0];
        __syncthreads();
        wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, g.N(N_ATdy), g.Q.blk, g.Q.nblk,
                EpiMnQATdy{nullptr, g.N(N_QATdy), g.N(N_dQw), g.N(N_ATdy), g.N(N_fact_M), diag}, sums, maxs, g.prod);
        wg_reduce<3, 0>(sums, maxs, g.red);
        const double dqw_atdy = g.red[0], atdy_qatdy = g.red[1], fm_term = g.red[2];
        __syncthreads();
        wg_map(m, MapDot{nullptr, g.M(M_dy), g.M(M_dy)}, sums, maxs);
        wg_reduce<1, 0>(sums, maxs, g.red);
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
__device__ double wg_golden_diag(WG &g, double a, double b)
{
    const int n = g.I->n;
    const double *dq = g.N(N_diag), *c = g.N(N_ATdy), *d = g.N(N_QATdy);
    auto eval = [&](double x) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < n; i += kThreads) acc += c[i] * (d[i] / (1.0 + x * dq[i]));
        const double tot = block_sum_all(acc);
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

__device__ void wg_update_sigma(WG &g, double *sums, double *maxs)
{
    State &S = *g.S;
    if (S.rs.flag < 1) return;
    const int n = g.I->n, m = g.I->m;
    const int diag = S.Q_is_diag;
    const double old = S.sigma;
    wg_map(n, MapSigmaDiffN{nullptr, g.N(N_dx), g.N(N_dw), g.N(N_ATdy), g.N(N_x_bar), g.N(N_last_x), g.N(N_w_bar),
                            g.N(N_last_w), g.N(N_ATy_bar), g.N(N_last_ATy), m > 0 ? 1 : 0},
           sums, maxs);
    wg_reduce<1, 0>(sums, maxs, g.red);
    double b = g.red[0];
    __syncthreads();
    double dy_sq = 0.0;
    if (m > 0) {
        wg_map(m, MapSigmaDiffM{nullptr, g.M(M_dy), g.M(M_y_bar), g.M(M_last_y)}, sums, maxs);
        wg_reduce<1, 0>(sums, maxs, g.red);
        dy_sq = g.red[0];
        __syncthreads();
    }
    wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, g.N(N_dw), g.Q.blk, g.Q.nblk,
            EpiSigmaQdw{nullptr, g.N(N_dQw), g.N(N_ATdy), g.N(N_dw)}, sums, maxs, g.prod);
    wg_reduce<3, 0>(sums, maxs, g.red);
    const double dqw_atdy = g.red[0], dqw_sq = g.red[1], dw_dqw = g.red[2];
    __syncthreads();
    double atdy_qatdy = 0.0;
    if (m > 0) {
        wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, g.N(N_ATdy), g.Q.blk, g.Q.nblk,
                EpiSigmaQATdy{nullptr, g.N(N_QATdy), g.N(N_ATdy)}, sums, maxs, g.prod);
        wg_reduce<1, 0>(sums, maxs, g.red);
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
        wg_map(n, MapDiagFactors{{}, g.N(N_fact1), g.N(N_fact2), g.N(N_fact), g.N(N_fact_M), g.N(N_diag), sg}, sums,
               maxs);
    __syncthreads();
}

// control.c:413-457
__device__ void wg_restart_apply(WG &g, double *sums, double *maxs)
{
    State &S = *g.S;
    if (S.rs.flag <= 0) return;
    const int m = g.I->m;
    wg_spmv(g.AT.rp, g.AT.ci, g.AT.val, g.M(M_y_bar), g.AT.blk, g.AT.nblk,
            EpiRestart{{}, g.N(N_ATy_bar), g.N(N_ATy), g.N(N_last_ATy), g.N(N_x), g.N(N_w), g.N(N_last_x),
                       g.N(N_last_w), g.N(N_Qw), g.N(N_last_Qw), g.N(N_x_bar), g.N(N_w_bar), g.N(N_Qw_bar),
                       m > 0 ? 1 : 0, g.I->noC},
            sums, maxs, g.prod);
    wg_map(m, MapRestartM{{}, g.M(M_y), g.M(M_last_y), g.M(M_y_bar)}, sums, maxs);
    if (threadIdx.x == 0) {
        S.rs.last_gap = S.rs.current_gap;
        S.rs.save_gap = INFINITY;
        S.rs.times += 1;
        S.rs.inner = 0;
    }
    __syncthreads();
}

// iteration.c:250-293 (general variant)
template <bool DIAG, bool CHECK>
__device__ void wg_step_t(WG &g, double h1, double h2, double *sums, double *maxs)
{
    const State &S = *g.S;
    const int m = g.I->m;
    const double sg = S.sigma;
    const double f2s = 1.0 / (1.0 + sg * S.lam_q), f1s = 1.0 - f2s;
    wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, g.N(N_w), g.Q.blk, g.Q.nblk,
            EpiZxw1<DIAG, CHECK>{{}, g.N(N_Qw), g.N(N_x), g.N(N_x_bar), g.N(N_x_hat), g.N(N_w_bar), g.N(N_dx),
                                 g.N(N_z_bar), g.N(N_w), g.N(N_ATy), g.N(N_c), g.N(N_l), g.N(N_u), g.N(N_last_x),
                                 g.N(N_fact1), g.N(N_fact2), sg, f1s, f2s, h1, h2},
            sums, maxs, g.prod);
    if (m > 0) {
        wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, g.N(N_w_bar), g.Q.blk, g.Q.nblk,
                EpiTempv{{}, g.N(N_tempv), g.N(N_x_hat), g.N(N_Qw), sg}, sums, maxs, g.prod);
        const double f1 = S.lam_a * sg, f2 = 1.0 / f1;
        wg_spmv(g.A.rp, g.A.ci, g.A.val, g.N(N_tempv), g.A.blk, g.A.nblk,
                EpiY<CHECK>{{}, g.M(M_y), g.M(M_y_bar), g.M(M_s), g.M(M_dy), g.M(M_AL), g.M(M_AU), g.M(M_last_y), f1,
                            f2, h1, h2},
                sums, maxs, g.prod);
    }
    const double fs = sg / (1.0 + sg * S.lam_q);
    wg_spmv(g.AT.rp, g.AT.ci, g.AT.val, g.M(M_y_bar), g.AT.blk, g.AT.nblk,
            EpiW2<DIAG, CHECK>{{}, g.N(N_w_bar), g.N(N_w), g.N(N_ATy), g.N(N_ATy_bar), g.N(N_dw), g.N(N_ATdy),
                               g.N(N_last_w), g.N(N_last_ATy), g.N(N_fact), fs, h1, h2},
            sums, maxs, g.prod);
}

// iteration.c:19-83 and 176-216 (noC: A^T y first, Halpern on w and Qw)
template <bool DIAG, bool CHECK>
__device__ void wg_step_noc_t(WG &g, double h1, double h2, double *sums, double *maxs)
{
    const State &S = *g.S;
    const int m = g.I->m;
    const double sg = S.sigma;
    const double f2s = 1.0 / (1.0 + sg * S.lam_q), f1s = 1.0 - f2s;
    wg_spmv(g.AT.rp, g.AT.ci, g.AT.val, g.M(M_y), g.AT.blk, g.AT.nblk,
            EpiZxw1NoC<DIAG, CHECK>{{}, g.N(N_ATy), g.N(N_x), g.N(N_x_bar), g.N(N_x_hat), g.N(N_w), g.N(N_w_bar),
                                    g.N(N_dx), g.N(N_dw), g.N(N_z_bar), g.N(N_Qw), g.N(N_c), g.N(N_l), g.N(N_u),
                                    g.N(N_last_x), g.N(N_last_w), g.N(N_fact1), g.N(N_fact2), sg, f1s, f2s, h1, h2},
            sums, maxs, g.prod);
    wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, g.N(N_w_bar), g.Q.blk, g.Q.nblk,
            EpiQwNoC{{}, g.N(N_Qw), g.N(N_Qw_bar), g.N(N_tempv), g.N(N_last_Qw), g.N(N_x_hat), sg, h1, h2}, sums,
            maxs, g.prod);
    if (m > 0) {
        const double f1 = S.lam_a * sg, f2 = 1.0 / f1;
        wg_spmv(g.A.rp, g.A.ci, g.A.val, g.N(N_tempv), g.A.blk, g.A.nblk,
                EpiY<CHECK>{{}, g.M(M_y), g.M(M_y_bar), g.M(M_s), g.M(M_dy), g.M(M_AL), g.M(M_AU), g.M(M_last_y), f1,
                            f2, h1, h2},
                sums, maxs, g.prod);
    }
}

__device__ void wg_step(WG &g, long long inner, double *sums, double *maxs)
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
__device__ bool wg_body(WG &g, double *sums, double *maxs)
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

__device__ void wg_bind(WG &g, const Arrays &ar, const Params &P, int k, State *S, double *prod, double *red)
{
    const Inst *I = ar.inst + k;
    g.I = I;
    g.S = S;
    g.P = P;
    g.prod = prod;
    g.red = red;
    g.Q = Mat{ar.rp + I->q_rp, ar.ci + I->q_nz, ar.val + I->q_nz, ar.blocks + I->q_blk, I->q_nblk};
    g.A = Mat{ar.rp + I->a_rp, ar.ci + ar.a_base + I->a_nz, ar.val + ar.a_base + I->a_nz, ar.blocks + I->a_blk,
              I->a_nblk};
    g.AT = Mat{ar.rp + I->at_rp, ar.ci + ar.at_ci_base + I->a_nz, ar.val + ar.at_ci_base + I->a_nz,
               ar.blocks + I->at_blk, I->at_nblk};
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

__device__ void adj_op(WG &g, const AdjMask &mk, const double *xn, const double *xm, double *yn, double *ym,
                       double *sums, double *maxs)
{
    double *tmp = g.N(N_pt);
    wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, xn, g.Q.blk, g.Q.nblk, EpiStore{{}, tmp}, sums, maxs, g.prod);
    wg_spmv(g.AT.rp, g.AT.ci, g.AT.val, xm, g.AT.blk, g.AT.nblk, EpiAdjN{{}, yn, tmp, mk}, sums, maxs, g.prod);
    if (g.I->m > 0) wg_spmv(g.A.rp, g.A.ci, g.A.val, xn, g.A.blk, g.A.nblk, EpiAdjM{{}, ym, mk}, sums, maxs, g.prod);
}

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

// y += a x over both parts
__device__ void adj_axpy(WG &g, double a, const double *xn, const double *xm, double *yn, double *ym)
{
    const int n = g.I->n, m = g.I->m;
    for (int i = threadIdx.x; i < n; i += kThreads) yn[i] += a * xn[i];
    for (int i = threadIdx.x; i < m; i += kThreads) ym[i] += a * xm[i];
    __syncthreads();
}

__device__ void adj_copy(WG &g, const double *xn, const double *xm, double *yn, double *ym)
{
    const int n = g.I->n, m = g.I->m;
    for (int i = threadIdx.x; i < n; i += kThreads) yn[i] = xn[i];
    for (int i = threadIdx.x; i < m; i += kThreads) ym[i] = xm[i];
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void batch_adjoint_kernel(Arrays ar, Params P, int count,
                                                                 const double *__restrict__ gvec,
                                                                 double *__restrict__ u_out,
                                                                 double *__restrict__ vrow_out,
                                                                 double *__restrict__ vbnd_out, double tol,
                                                                 int max_steps, int *__restrict__ steps_out)
{
    __shared__ double prod[kStreamCap];
    __shared__ double red[kMaxTerms];
    __shared__ State S;
    const int k = blockIdx.x;
    if (k >= count) return;
    if (threadIdx.x == 0) S = ar.state[k];
    __syncthreads();
    WG g;
    wg_bind(g, ar, P, k, &S, prod, red);
    const int n = g.I->n, m = g.I->m, t = threadIdx.x;
    double sums[kMaxTerms] = {0}, maxs[kMaxTerms] = {0};
    const AdjMask mk{g.N(N_z_bar), g.N(N_l), g.N(N_u), g.M(M_y_bar), g.M(M_AL), g.M(M_AU)};
    const double *d = g.N(N_col), *rr = g.M(M_row);
    const double *gk = gvec + g.I->vn;
    double *bn = g.N(N_tc);
    double *xn = g.N(N_pq), *xm = g.M(M_pq);
    double *rn = g.N(N_pz), *rm = g.M(M_pz);
    double *pn[3] = {g.N(N_dx), g.N(N_dw), g.N(N_tempv)};
    double *pm[3] = {g.M(M_tr), g.M(M_dy), g.M(M_e0)};
    double *sn[3] = {g.N(N_dQw), g.N(N_QATdy), g.N(N_Qx)};
    double *sm[3] = {g.M(M_e1), g.M(M_e2), g.M(M_e3)};
    for (int i = t; i < n; i += kThreads) {
        const double b = mk.free_var(i) ? gk[i] / d[i] : 0.0;
        bn[i] = b;
        xn[i] = 0.0;
        rn[i] = b;
        pn[0][i] = b;
    }
    for (int i = t; i < m; i += kThreads) {
        xm[i] = 0.0;
        rm[i] = 0.0;
        pm[0][i] = 0.0;
    }
    __syncthreads();
    const double bnorm = sqrt(adj_dot(g, bn, rm, bn, rm, sums, maxs));
    adj_op(g, mk, pn[0], pm[0], sn[0], sm[0], sums, maxs);
    adj_copy(g, pn[0], pm[0], pn[1], pm[1]);
    adj_copy(g, sn[0], sm[0], sn[1], sm[1]);
    // ring indices: i0 = newest (p0,s0), i1 = p1,s1, i2 = p2,s2
    int i0 = 0, i1 = 1, i2 = 2;
    int step = 0;
    if (bnorm > 0.0) {
        for (step = 1; step <= max_steps; ++step) {
            // rotate: p2 <- p1, p1 <- p0; the next p0 is written into p2's old buffer
            const int o2 = i2;
            i2 = i1;
            i1 = i0;
            i0 = o2;
            const double ss = adj_dot(g, sn[i1], sm[i1], sn[i1], sm[i1], sums, maxs);
            if (!(ss > 0.0)) break;
            const double alpha = adj_dot(g, rn, rm, sn[i1], sm[i1], sums, maxs) / ss;
            adj_axpy(g, alpha, pn[i1], pm[i1], xn, xm);
            adj_axpy(g, -alpha, sn[i1], sm[i1], rn, rm);
            const double rnorm = sqrt(adj_dot(g, rn, rm, rn, rm, sums, maxs));
            if (rnorm <= tol * bnorm) break;
            adj_copy(g, sn[i1], sm[i1], pn[i0], pm[i0]);
            adj_op(g, mk, sn[i1], sm[i1], sn[i0], sm[i0], sums, maxs);
            const double beta1 = adj_dot(g, sn[i0], sm[i0], sn[i1], sm[i1], sums, maxs) / ss;
            adj_axpy(g, -beta1, pn[i1], pm[i1], pn[i0], pm[i0]);
            adj_axpy(g, -beta1, sn[i1], sm[i1], sn[i0], sm[i0]);
            if (step > 1) {
                const double s22 = adj_dot(g, sn[i2], sm[i2], sn[i2], sm[i2], sums, maxs);
                if (s22 > 0.0) {
                    const double beta2 = adj_dot(g, sn[i0], sm[i0], sn[i2], sm[i2], sums, maxs) / s22;
                    adj_axpy(g, -beta2, pn[i2], pm[i2], pn[i0], pm[i0]);
                    adj_axpy(g, -beta2, sn[i2], sm[i2], sn[i0], sm[i0]);
                }
            }
        }
    }
    // bound multipliers: v~_B = b~_B - (Q~ u~ + A~' v~)_B with b~ = g/d
    double *tn = g.N(N_Qw_bar), *tmp = g.N(N_pt);
    wg_spmv(g.Q.rp, g.Q.ci, g.Q.val, xn, g.Q.blk, g.Q.nblk, EpiStore{{}, tmp}, sums, maxs, g.prod);
    wg_spmv(g.AT.rp, g.AT.ci, g.AT.val, xm, g.AT.blk, g.AT.nblk, EpiAddStore{{}, tn, tmp}, sums, maxs, g.prod);
    for (int i = t; i < n; i += kThreads) {
        const bool fr = mk.free_var(i);
        u_out[g.I->vn + i] = xn[i] / d[i];
        vbnd_out[g.I->vn + i] = fr ? 0.0 : d[i] * (gk[i] / d[i] - tn[i]);
    }
    for (int i = t; i < m; i += kThreads) vrow_out[g.I->vm + i] = xm[i] / rr[i];
    if (t == 0) steps_out[k] = step;
}

__global__ __launch_bounds__(kThreads) void batch_setup_kernel(Arrays ar, Params P, int count)
{
    __shared__ double prod[kStreamCap];
    __shared__ double red[kMaxTerms];
    __shared__ State S;
    const int k = blockIdx.x;
    if (k >= count) return;
    WG g;
    wg_bind(g, ar, P, k, &S, prod, red);
    if (threadIdx.x == 0) {
        S = State{};
        S.first4 = S.first6 = 1;
        S.kkt = INFINITY;
        S.rs.first = 1;
        S.rs.save_gap = S.rs.current_gap = S.rs.last_gap = S.rs.best_gap = INFINITY;
        S.rs.weighted_norm = INFINITY;
        S.rs.best_sigma = 1.0;
        S.status = HPRQP_STATUS_CONTINUE;
    }
    __syncthreads();
    double sums[kMaxTerms] = {0}, maxs[kMaxTerms] = {0};
    wg_setup(g, sums, maxs);
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
__global__ __launch_bounds__(kThreads) void batch_solve_kernel(Arrays ar, Params P, int count, int bodies)
{
    __shared__ double prod[kStreamCap];
    __shared__ double red[kMaxTerms];
    __shared__ State S;
    const int k = blockIdx.x;
    if (k >= count) return;
    if (threadIdx.x == 0) S = ar.state[k];
    __syncthreads();
    if (S.finished) return;
    WG g;
    wg_bind(g, ar, P, k, &S, prod, red);
    double sums[kMaxTerms] = {0}, maxs[kMaxTerms] = {0};
    for (int b = 0; b < bodies; ++b)
        if (wg_body(g, sums, maxs)) break;
    if (threadIdx.x == 0) {
        ar.state[k] = S;
        if (!S.finished) atomicAdd(ar.active, 1);
    }
}

}  // namespace batch
}  // namespace hprqp

// ================================================================ host side

struct HPRQPHipBatch {
    size_t count = 0;
    hprqp::batch::Arrays ar{};
    hprqp::batch::Params P{};
    std::vector<hprqp::batch::Inst> inst;
    std::vector<size_t> n, m;
    hipStream_t stream = nullptr;
    double *d_vn = nullptr, *d_vm = nullptr, *d_val = nullptr;
    int *d_rp = nullptr, *d_ci = nullptr, *d_perm = nullptr, *d_blocks = nullptr, *d_active = nullptr;
    hprqp::batch::Inst *d_inst = nullptr;
    hprqp::batch::State *d_state = nullptr;
    std::vector<hprqp::batch::State> state;
    double time_limit = 3600.0;
    double solve_seconds = 0.0;
    size_t launches = 0;
    ~HPRQPHipBatch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : {(void *)d_vn, (void *)d_vm, (void *)d_val, (void *)d_rp, (void *)d_ci, (void *)d_perm,
                        (void *)d_blocks, (void *)d_active, (void *)d_inst, (void *)d_state})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Row blocks for the workgroup engine: runs of rows with >= kWaveRow entries
// (up to 8 per block, one wavefront per row), otherwise stream blocks of
// <= kThreads rows and <= kStreamCap entries.
std::vector<int> plan_batch_blocks(const std::vector<int> &rp)
{
    using hprqp::batch::kWaveRow;
    const int nrow = (int)rp.size() - 1;
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
            if (len >= kWaveRow || nnz + len > hprqp::kStreamCap) break;
            nnz += len;
            ++r;
        }
    }
    starts.push_back(nrow);
    if (nrow == 0) starts = {0, 0};
    return starts;
}

// Host counting-sort transpose permutation (column-major order, rows
// ascending within a column): perm[k] = source entry of output entry k.
void transpose_perm(size_t nrow, size_t ncol, const size_t *rp, const size_t *ci, std::vector<int> &trp,
                    std::vector<int> &tci, std::vector<int> &perm)
{
    const size_t nnz = rp[nrow];
    trp.assign(ncol + 1, 0);
    tci.resize(nnz);
    perm.resize(nnz);
    for (size_t k = 0; k < nnz; ++k) trp[ci[k] + 1] += 1;
    for (size_t j = 0; j < ncol; ++j) trp[j + 1] += trp[j];
    std::vector<int> fill(trp.begin(), trp.end() - 1);
    for (size_t r = 0; r < nrow; ++r)
        for (size_t k = rp[r]; k < rp[r + 1]; ++k) {
            const int dst = fill[ci[k]]++;
            tci[dst] = (int)r;
            perm[dst] = (int)k;
        }
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
    HPRQP_HIP(hipSetDevice(params->device_number));
    auto *b = new (std::nothrow) HPRQPHipBatch();
    if (!b) return HPRQP_STATUS_ALLOCATION_FAILED;
    b->count = count;
    b->time_limit = params->time_limit;
    b->P = Params{params->stoptol, (long long)std::min(params->max_iter, (size_t)INT64_MAX),
                  (long long)params->check_iter, params->print_frequency, params->use_scaling};
    // sizes and validation
    long long sum_n = 0, sum_m = 0, sum_q = 0, sum_a = 0, sum_rp = 0;
    for (size_t k = 0; k < count; ++k) {
        const HPRQPModel &md = models[k];
        if (md.num_var_cones || md.num_row_cones || md.Q.ncol == 0 || !matrix_csr_is_valid(&md.Q) ||
            !matrix_csr_is_valid(&md.A)) {
            delete b;
            return md.num_var_cones || md.num_row_cones ? HPRQP_STATUS_NOT_IMPLEMENTED : HPRQP_STATUS_INVALID_ARGUMENT;
        }
        sum_n += (long long)md.Q.ncol;
        sum_m += (long long)md.A.nrow;
        sum_q += (long long)md.Q.nnz;
        sum_a += (long long)md.A.nnz;
        sum_rp += 3 * ((long long)md.Q.ncol + 1) + (long long)md.A.nrow + 1;
    }
    if (sum_q + 2 * sum_a > INT32_MAX || sum_rp > INT32_MAX) {
        delete b;
        return HPRQP_STATUS_INVALID_ARGUMENT;
    }
    std::vector<int> rp, ci, perm, blocks;
    std::vector<double> val;
    rp.reserve(sum_rp);
    ci.resize(sum_q + 2 * sum_a);
    val.resize(sum_q + 2 * sum_a, 0.0);
    perm.resize(sum_a + sum_q);
    std::vector<double> vn((size_t)sum_n * N_COUNT, 0.0), vm((size_t)std::max(sum_m, 1LL) * M_COUNT, 0.0);
    b->inst.resize(count);
    long long qnz = 0, anz = 0, vno = 0, vmo = 0;
    const long long a_base = sum_q, at_base = sum_q + sum_a;
    std::vector<int> trp, tci, tperm;
    for (size_t k = 0; k < count; ++k) {
        const HPRQPModel &md = models[k];
        Inst &I = b->inst[k];
        const size_t n = md.Q.ncol, m = md.A.nrow;
        I.n = (int)n;
        I.m = (int)m;
        I.q_nz = qnz;
        I.a_nz = anz;
        I.vn = vno;
        I.vm = vmo;
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
        auto push_rp = [&](const std::vector<int> &v) {
            long long off = (long long)rp.size();
            rp.insert(rp.end(), v.begin(), v.end());
            return off;
        };
        auto push_blocks = [&](const std::vector<int> &r, int &off, int &cnt) {
            std::vector<int> blk = plan_batch_blocks(r);
            off = (int)blocks.size();
            cnt = (int)blk.size() - 1;
            blocks.insert(blocks.end(), blk.begin(), blk.end());
        };
        std::vector<int> qrp(n + 1), arp(m + 1);
        for (size_t r = 0; r <= n; ++r) qrp[r] = (int)md.Q.row_ptr[r];
        for (size_t r = 0; r <= m; ++r) arp[r] = (int)md.A.row_ptr[r];
        I.q_rp = push_rp(qrp);
        I.a_rp = push_rp(arp);
        for (size_t e = 0; e < md.Q.nnz; ++e) {
            ci[qnz + e] = (int)md.Q.col_ind[e];
            val[qnz + e] = md.Q.values[e];
        }
        for (size_t e = 0; e < md.A.nnz; ++e) {
            ci[a_base + anz + e] = (int)md.A.col_ind[e];
            val[a_base + anz + e] = md.A.values[e];
        }
        transpose_perm(m, n, md.A.row_ptr, md.A.col_ind, trp, tci, tperm);
        I.at_rp = push_rp(trp);
        for (size_t e = 0; e < md.A.nnz; ++e) {
            ci[at_base + anz + e] = tci[e];
            perm[anz + e] = tperm[e];
        }
        transpose_perm(n, n, md.Q.row_ptr, md.Q.col_ind, trp, tci, tperm);
        I.qt_rp = push_rp(trp);
        for (size_t e = 0; e < md.Q.nnz; ++e) perm[sum_a + qnz + e] = tperm[e];
        push_blocks(qrp, I.q_blk, I.q_nblk);
        push_blocks(arp, I.a_blk, I.a_nblk);
        std::vector<int> atrp(rp.begin() + I.at_rp, rp.begin() + I.at_rp + n + 1);
        push_blocks(atrp, I.at_blk, I.at_nblk);
        for (size_t i = 0; i < n; ++i) {
            vn[(size_t)N_c * sum_n + vno + i] = md.c[i];
            vn[(size_t)N_l * sum_n + vno + i] = md.l[i];
            vn[(size_t)N_u * sum_n + vno + i] = md.u[i];
        }
        for (size_t i = 0; i < m; ++i) {
            vm[(size_t)M_AL * sum_m + vmo + i] = md.AL[i];
            vm[(size_t)M_AU * sum_m + vmo + i] = md.AU[i];
        }
        b->n.push_back(n);
        b->m.push_back(m);
        qnz += (long long)md.Q.nnz;
        anz += (long long)md.A.nnz;
        vno += (long long)n;
        vmo += (long long)m;
    }
    HPRQPStatus st = HPRQP_STATUS_OK;
    auto upload = [&](auto *&dst, const auto &src) -> bool {
        using T = typename std::remove_reference<decltype(src)>::type::value_type;
        size_t bytes = sizeof(T) * std::max(src.size(), (size_t)1);
        if (hipMalloc((void **)&dst, bytes) != hipSuccess) return false;
        if (!src.empty() && hipMemcpy(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice) != hipSuccess)
            return false;
        return true;
    };
    if (!upload(b->d_rp, rp) || !upload(b->d_ci, ci) || !upload(b->d_val, val) || !upload(b->d_perm, perm) ||
        !upload(b->d_blocks, blocks) || !upload(b->d_vn, vn) || !upload(b->d_vm, vm) || !upload(b->d_inst, b->inst))
        st = HPRQP_STATUS_ALLOCATION_FAILED;
    if (st == HPRQP_STATUS_OK) {
        if (hipMalloc((void **)&b->d_state, sizeof(State) * count) != hipSuccess ||
            hipMalloc((void **)&b->d_active, sizeof(int)) != hipSuccess ||
            hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
            st = HPRQP_STATUS_ALLOCATION_FAILED;
    }
    if (st != HPRQP_STATUS_OK) {
        delete b;
        return st;
    }
    Arrays &ar = b->ar;
    ar.inst = b->d_inst;
    ar.state = b->d_state;
    ar.rp = b->d_rp;
    ar.ci = b->d_ci;
    ar.val = b->d_val;
    ar.perm = b->d_perm;
    ar.blocks = b->d_blocks;
    ar.vn = b->d_vn;
    ar.vm = b->d_vm;
    ar.sum_n = sum_n;
    ar.sum_m = std::max(sum_m, 1LL);
    ar.at_ci_base = at_base;
    ar.a_base = a_base;
    ar.qt_perm_base = sum_a;
    ar.active = b->d_active;
    batch_setup_kernel<<<(unsigned)count, hprqp::kThreads, 0, b->stream>>>(ar, b->P, (int)count);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) {
        delete b;
        return HPRQP_STATUS_NOT_IMPLEMENTED;
    }
    *out = b;
    return HPRQP_STATUS_OK;
}

// Relaunches the chunk kernel (bodies_per_launch loop bodies per instance)
// until every instance finished or the time limit passed.
HPRQPStatus hprqp_hip_batch_solve(HPRQPHipBatch *b, size_t bodies_per_launch, size_t *remaining)
{
    using namespace hprqp::batch;
    if (!b || bodies_per_launch == 0) return HPRQP_STATUS_INVALID_ARGUMENT;
    const double t0 = hprqp::wall_seconds();
    int active = 1;
    while (active > 0) {
        HPRQP_HIP(hipMemsetAsync(b->d_active, 0, sizeof(int), b->stream));
        batch_solve_kernel<<<(unsigned)b->count, hprqp::kThreads, 0, b->stream>>>(b->ar, b->P, (int)b->count,
                                                                                  (int)bodies_per_launch);
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

void hprqp_hip_batch_destroy(HPRQPHipBatch *b) { delete b; }

}  // extern "C"

extern "C" {

// Adjoint of every instance of a solved batch: for upstream gradients g
// (dL/dx in original units, instances concatenated) returns u (Sum n), the
// row multipliers v_rows (Sum m) and the bound multipliers v_bounds (Sum n) of
// the active-set KKT system K [u; v] = [g; 0]; steps[k] = MINRES steps used.
// Host arrays in and out.
HPRQPStatus hprqp_hip_batch_adjoint(HPRQPHipBatch *b, const double *dl_dx, double tol, int max_steps, double *u,
                                    double *v_rows, double *v_bounds, int *steps)
{
    using namespace hprqp::batch;
    if (!b || !dl_dx || !u || !v_rows || !v_bounds || max_steps <= 0) return HPRQP_STATUS_INVALID_ARGUMENT;
    const size_t sn = (size_t)b->ar.sum_n, sm = (size_t)b->ar.sum_m;
    double *d_g = nullptr, *d_u = nullptr, *d_vr = nullptr, *d_vb = nullptr;
    int *d_steps = nullptr;
    HPRQP_HIP(hipMalloc((void **)&d_g, sizeof(double) * sn));
    HPRQP_HIP(hipMalloc((void **)&d_u, sizeof(double) * sn));
    HPRQP_HIP(hipMalloc((void **)&d_vb, sizeof(double) * sn));
    HPRQP_HIP(hipMalloc((void **)&d_vr, sizeof(double) * sm));
    HPRQP_HIP(hipMalloc((void **)&d_steps, sizeof(int) * b->count));
    HPRQP_HIP(hipMemcpy(d_g, dl_dx, sizeof(double) * sn, hipMemcpyHostToDevice));
    HPRQP_HIP(hipMemset(d_vr, 0, sizeof(double) * sm));
    batch_adjoint_kernel<<<(unsigned)b->count, hprqp::kThreads, 0, b->stream>>>(b->ar, b->P, (int)b->count, d_g, d_u,
                                                                                d_vr, d_vb, tol, max_steps, d_steps);
    HPRQP_HIP(hipGetLastError());
    HPRQP_HIP(hipStreamSynchronize(b->stream));
    HPRQP_HIP(hipMemcpy(u, d_u, sizeof(double) * sn, hipMemcpyDeviceToHost));
    HPRQP_HIP(hipMemcpy(v_bounds, d_vb, sizeof(double) * sn, hipMemcpyDeviceToHost));
    size_t real_m = 0;
    for (size_t k = 0; k < b->count; ++k) real_m += b->m[k];
    HPRQP_HIP(hipMemcpy(v_rows, d_vr, sizeof(double) * real_m, hipMemcpyDeviceToHost));
    if (steps) HPRQP_HIP(hipMemcpy(steps, d_steps, sizeof(int) * b->count, hipMemcpyDeviceToHost));
    (void)hipFree(d_g);
    (void)hipFree(d_u);
    (void)hipFree(d_vr);
    (void)hipFree(d_vb);
    (void)hipFree(d_steps);
    return HPRQP_STATUS_OK;
}

}  // extern "C"
