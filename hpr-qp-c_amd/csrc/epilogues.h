// Per-row epilogues of the fused SpMV kernels and the elementwise map
// functors of the HPR iteration. Each mirrors one loop of the reference's
// CPU path (src/cpu/iteration.c, residuals.c, control.c, eigs.c,
// workspace.c) with the same arithmetic order; kernels that use them are
// compiled with -ffp-contract=off. Shared by the single-QP backend and the
// batched (workgroup-per-instance) solver.
#pragma once

#include <hip/hip_runtime.h>

namespace hprqp {

// iteration.c:8-17
__device__ __forceinline__ double project(double v, double lo, double hi)
{
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}


struct NoRed {
    static constexpr int NS = 0, NM = 0;
    double *partials = nullptr;
};

struct EpiStore : NoRed {
    double *out;
    __device__ void row(int i, double v, double *, double *) const { out[i] = v; }
};

// Partial row sum handed to the next column panel.
struct EpiCarry : NoRed {
    double *carry;
    __device__ void row(int i, double v, double *, double *) const { carry[i] = v; }
};

// iteration.c:85-123
template <bool DIAG, bool CHECK>
struct EpiZxw1 : NoRed {
    double *Qw, *x, *x_bar, *x_hat, *w_bar, *dx, *z_bar;
    const double *w, *ATy, *c, *l, *u, *last_x, *fact1, *fact2;
    double sigma, f1s, f2s, h1, h2;
    // row inputs, loaded before the product is known
    struct Pre {
        double xi, aty, ci, li, ui, lx, wi, f1, f2;
    };
    __device__ Pre load(int i) const
    {
        Pre p;
        p.xi = x[i];
        p.aty = ATy[i];
        p.ci = c[i];
        p.li = l[i];
        p.ui = u[i];
        p.lx = last_x[i];
        p.wi = w[i];
        p.f1 = DIAG ? fact1[i] : f1s;
        p.f2 = DIAG ? fact2[i] : f2s;
        return p;
    }
    __device__ void row_pre(int i, double qw, const Pre &p, double *, double *) const
    {
        Qw[i] = qw;
        const double tmp = -qw + p.aty - p.ci;
        const double zr = p.xi + sigma * tmp;
        const double xb = project(zr, p.li, p.ui);
        const double xh = 2.0 * xb - p.xi;
        const double xn = h2 * xh + h1 * p.lx;
        const double wb = p.f1 * p.wi + p.f2 * xh;
        if (CHECK) {
            dx[i] = xb - p.xi;
            z_bar[i] = (xb - zr) / sigma;
        }
        x_bar[i] = xb;
        x_hat[i] = xh;
        x[i] = xn;
        w_bar[i] = wb;
    }
    __device__ void row(int i, double qw, double *s, double *m) const { row_pre(i, qw, load(i), s, m); }
};

// iteration.c:141-143
struct EpiTempv : NoRed {
    double *tempv;
    const double *x_hat, *Qw;
    double sigma;
    struct Pre {
        double xh, qw;
    };
    __device__ Pre load(int i) const { return Pre{x_hat[i], Qw[i]}; }
    __device__ void row_pre(int i, double qwb, const Pre &p, double *, double *) const
    {
        tempv[i] = p.xh + sigma * (p.qw - qwb);
    }
    __device__ void row(int i, double qwb, double *s, double *m) const { row_pre(i, qwb, load(i), s, m); }
};

// iteration.c:150-171
template <bool CHECK>
struct EpiY : NoRed {
    double *y, *y_bar, *s, *dy;
    const double *AL, *AU, *last_y;
    double f1, f2, h1, h2;
    struct Pre {
        double yi, al, au, ly;
    };
    __device__ Pre load(int i) const { return Pre{y[i], AL[i], AU[i], last_y[i]}; }
    __device__ void row_pre(int i, double ax, const Pre &p, double *, double *) const
    {
        const double sr = ax - f1 * p.yi;
        const double sp = project(sr, p.al, p.au);
        const double yb = f2 * (sp - sr);
        const double yh = 2.0 * yb - p.yi;
        const double yn = h2 * yh + h1 * p.ly;
        if (CHECK) {
            s[i] = sp;
            dy[i] = yb - p.yi;
        }
        y_bar[i] = yb;
        y[i] = yn;
    }
    __device__ void row(int i, double ax, double *sm, double *m) const { row_pre(i, ax, load(i), sm, m); }
};

// iteration.c:218-248
template <bool DIAG, bool CHECK>
struct EpiW2 : NoRed {
    double *w_bar, *w, *ATy, *ATy_bar, *dw, *ATdy;
    const double *last_w, *last_ATy, *fact;
    double fs, h1, h2;
    struct Pre {
        double f, aty, wi, wb, lw, laty;
    };
    __device__ Pre load(int i) const
    {
        return Pre{DIAG ? fact[i] : fs, ATy[i], w[i], w_bar[i], last_w[i], last_ATy[i]};
    }
    __device__ void row_pre(int i, double atyb, const Pre &p, double *, double *) const
    {
        const double wbn = p.wb + p.f * (atyb - p.aty);
        const double wn = h1 * p.lw + h2 * (2.0 * wbn - p.wi);
        const double atyn = h1 * p.laty + h2 * (2.0 * atyb - p.aty);
        if (CHECK) {
            dw[i] = wbn - p.wi;
            ATdy[i] = atyb - p.aty;
            ATy_bar[i] = atyb;
        }
        w_bar[i] = wbn;
        w[i] = wn;
        ATy[i] = atyn;
    }
    __device__ void row(int i, double atyb, double *s, double *m) const { row_pre(i, atyb, load(i), s, m); }
};

// iteration.c:19-66 (noC: A^T y first, Halpern on w as well)
template <bool DIAG, bool CHECK>
struct EpiZxw1NoC : NoRed {
    double *ATy, *x, *x_bar, *x_hat, *w, *w_bar, *dx, *dw, *z_bar;
    const double *Qw, *c, *l, *u, *last_x, *last_w, *fact1, *fact2;
    double sigma, f1s, f2s, h1, h2;
    __device__ void row(int i, double aty, double *, double *) const
    {
        ATy[i] = aty;
        const double xi = x[i], wi = w[i];
        const double tmp = -Qw[i] + aty - c[i];
        const double zr = xi + sigma * tmp;
        const double xb = project(zr, l[i], u[i]);
        const double xh = 2.0 * xb - xi;
        const double xn = h2 * xh + h1 * last_x[i];
        const double wb = DIAG ? fact1[i] * wi + fact2[i] * xh : f1s * wi + f2s * xh;
        const double wh = 2.0 * wb - wi;
        const double wn = h2 * wh + h1 * last_w[i];
        if (CHECK) {
            dx[i] = xb - xi;
            dw[i] = wb - wi;
            z_bar[i] = (xb - zr) / sigma;
        }
        x_bar[i] = xb;
        x_hat[i] = xh;
        x[i] = xn;
        w_bar[i] = wb;
        w[i] = wn;
    }
};

// iteration.c:68-80
struct EpiQwNoC : NoRed {
    double *Qw, *Qw_bar, *tempv;
    const double *last_Qw, *x_hat;
    double sigma, h1, h2;
    __device__ void row(int i, double qwb, double *, double *) const
    {
        const double qw = Qw[i];
        const double qwh = 2.0 * qwb - qw;
        Qw_bar[i] = qwb;
        Qw[i] = h2 * qwh + h1 * last_Qw[i];
        tempv[i] = x_hat[i] + sigma * (qw - qwb);
    }
};

// residuals.c:102-160: Qx = Q x_bar and the three objective dots.
struct EpiResQ {
    static constexpr int NS = 3, NM = 0;
    double *partials;
    double *Qx;
    const double *x_bar, *c, *z_bar;
    __device__ void row(int i, double qx, double *s, double *) const
    {
        Qx[i] = qx;
        const double xb = x_bar[i];
        s[0] += c[i] * xb;
        s[1] += xb * qx;
        s[2] += z_bar[i] * xb;
    }
};

// residuals.c:148-162: A^T y_bar and the dual residual.
struct EpiResAT {
    static constexpr int NS = 0, NM = 3;
    double *partials;
    double *ATdy;
    const double *Qx, *c, *z_bar, *col_norm;
    double c_scale;
    __device__ void row(int i, double atdy, double *, double *mx) const
    {
        ATdy[i] = atdy;
        const double f = col_norm[i] * c_scale;
        const double q = Qx[i];
        const double rd = (q + c[i] - atdy - z_bar[i]) * f;
        mx[0] = fmax(mx[0], fabs(rd));
        mx[1] = fmax(mx[1], fabs(q * f));
        mx[2] = fmax(mx[2], fabs(atdy * f));
    }
};

// EpiResAT when A^T y_bar is already at hand (the check variant of the
// iteration's last SpMV wrote it): the same terms from a stored vector.
struct MapResAT {
    static constexpr int NS = 0, NM = 3;
    double *partials;
    EpiResAT epi;
    const double *ATy_bar;
    __device__ void row(int i, double *s, double *mx) const { epi.row(i, ATy_bar[i], s, mx); }
};

// residuals.c:164-175 and the y_bar.s dot of :53
struct EpiResA {
    static constexpr int NS = 1, NM = 2;
    double *partials;
    double *Ax;
    const double *AL, *AU, *row_norm, *y_bar;
    double *s;
    double b_scale;
    int reproject;  // 1: s <- the projection of A x_bar (residuals.cu:28-50)
    __device__ void row(int i, double ax, double *sm, double *mx) const
    {
        Ax[i] = ax;
        const double p = project(ax, AL[i], AU[i]);
        const double f = row_norm[i] * b_scale;
        mx[0] = fmax(mx[0], fabs((p - ax) * f));
        mx[1] = fmax(mx[1], fabs(ax * f));
        if (reproject) s[i] = p;
        sm[0] += y_bar[i] * (reproject ? p : s[i]);
    }
};

// residuals.c:58-78
struct MapBoundViolation {
    static constexpr int NS = 0, NM = 1;
    double *partials;
    double *dx;
    const double *l, *u, *x_bar, *col_norm;
    double b_scale;
    __device__ void row(int i, double *, double *mx) const
    {
        const double lo = l[i] - x_bar[i], hi = x_bar[i] - u[i];
        double v = 0.0;
        if (lo > 0.0) v += lo;
        if (hi > 0.0) v += hi;
        const double d = v * (b_scale / col_norm[i]);
        dx[i] = d;
        mx[0] = fmax(mx[0], d);
    }
};

// control.c:174-209: the M-norm products.
struct EpiMnQdw {
    static constexpr int NS = 4, NM = 0;
    double *partials;
    double *dQw;
    const double *dx, *dw;
    __device__ void row(int i, double v, double *s, double *) const
    {
        dQw[i] = v;
        const double d = dx[i];
        s[0] += d * d;
        s[1] += v * d;
        s[2] += dw[i] * v;
        s[3] += v * v;
    }
};

struct EpiMnATdy {
    static constexpr int NS = 1, NM = 0;
    double *partials;
    double *ATdy;
    const double *dx;
    __device__ void row(int i, double v, double *s, double *) const
    {
        ATdy[i] = v;
        s[0] += v * dx[i];
    }
};

struct EpiMnQATdy {
    static constexpr int NS = 3, NM = 0;
    double *partials;
    double *QATdy;
    const double *dQw, *ATdy, *fact_M;
    int diag;
    __device__ void row(int i, double v, double *s, double *) const
    {
        QATdy[i] = v;
        const double a = ATdy[i];
        s[0] += dQw[i] * a;
        s[1] += a * v;
        if (diag) s[2] += (a * fact_M[i]) * v;
    }
};

struct MapDot {
    static constexpr int NS = 1, NM = 0;
    double *partials;
    const double *a, *b;
    __device__ void row(int i, double *s, double *) const { s[0] += a[i] * b[i]; }
};

// control.c:333-356: differences to the last restart point.
struct MapSigmaDiffN {
    static constexpr int NS = 1, NM = 0;
    double *partials;
    double *dx, *dw, *ATdy;
    const double *x_bar, *last_x, *w_bar, *last_w, *ATy_bar, *last_ATy;
    int has_rows;
    __device__ void row(int i, double *s, double *) const
    {
        const double d = x_bar[i] - last_x[i];
        dx[i] = d;
        dw[i] = w_bar[i] - last_w[i];
        if (has_rows) ATdy[i] = ATy_bar[i] - last_ATy[i];
        s[0] += d * d;
    }
};

struct MapSigmaDiffM {
    static constexpr int NS = 1, NM = 0;
    double *partials;
    double *dy;
    const double *y_bar, *last_y;
    __device__ void row(int i, double *s, double *) const
    {
        const double d = y_bar[i] - last_y[i];
        dy[i] = d;
        s[0] += d * d;
    }
};

struct EpiSigmaQdw {
    static constexpr int NS = 3, NM = 0;
    double *partials;
    double *dQw;
    const double *ATdy, *dw;
    __device__ void row(int i, double v, double *s, double *) const
    {
        dQw[i] = v;
        s[0] += v * ATdy[i];
        s[1] += v * v;
        s[2] += dw[i] * v;
    }
};

struct EpiSigmaQATdy {
    static constexpr int NS = 1, NM = 0;
    double *partials;
    double *QATdy;
    const double *ATdy;
    __device__ void row(int i, double v, double *s, double *) const
    {
        QATdy[i] = v;
        s[0] += ATdy[i] * v;
    }
};

// control.c:116-127
struct MapDiagFactors : NoRed {
    double *fact1, *fact2, *fact, *fact_M;
    const double *diag_Q;
    double sigma;
    __device__ void row(int i, double *, double *) const
    {
        const double inv = 1.0 / (1.0 + sigma * diag_Q[i]);
        fact2[i] = inv;
        fact[i] = sigma * inv;
        fact1[i] = sigma * diag_Q[i] * inv;
        fact_M[i] = sigma * sigma * inv;
    }
};

// control.c:413-457 (n side fused with the A^T y_bar product)
struct EpiRestart : NoRed {
    double *ATy_bar, *ATy, *last_ATy, *x, *w, *last_x, *last_w, *Qw, *last_Qw;
    const double *x_bar, *w_bar, *Qw_bar;
    int has_rows, noC;
    __device__ void row(int i, double v, double *, double *) const
    {
        const double xb = x_bar[i], wb = w_bar[i];
        x[i] = xb;
        w[i] = wb;
        last_x[i] = xb;
        last_w[i] = wb;
        if (has_rows) {
            ATy_bar[i] = v;
            last_ATy[i] = v;
            ATy[i] = v;
        }
        if (noC) {
            Qw[i] = Qw_bar[i];
            last_Qw[i] = Qw_bar[i];
        }
    }
};

struct MapRestartM : NoRed {
    double *y, *last_y;
    const double *y_bar;
    __device__ void row(int i, double *, double *) const
    {
        y[i] = y_bar[i];
        last_y[i] = y_bar[i];
    }
};

// eigs.c:11-153 pieces
struct MapProbe {
    static constexpr int NS = 1, NM = 0;
    double *partials;
    double *z;
    double inv_len1;
    double len1;
    __device__ void row(int i, double *s, double *) const
    {
        const double v = 1.0 + ((double)i + 1.0) / len1;
        z[i] = v;
        s[0] += v * v;
    }
};

struct MapScaleDiv : NoRed {
    double *q;
    const double *z;
    double d;
    __device__ void row(int i, double *, double *) const { q[i] = z[i] / d; }
};

// q = z / sqrt(||z||^2) with ||z||^2 read on the device (IEEE sqrt and
// division, so the same bits as dividing by the host's std::sqrt of it)
struct MapScaleDivNorm : NoRed {
    double *q;
    const double *z;
    const double *z2;
    __device__ void row(int i, double *, double *) const { q[i] = z[i] / sqrt(*z2); }
};

struct EpiPowerZ {
    static constexpr int NS = 2, NM = 0;
    double *partials;
    double *z;
    const double *q;
    __device__ void row(int i, double v, double *s, double *) const
    {
        z[i] = v;
        s[0] += q[i] * v;
        s[1] += v * v;
    }
};

// ||z - lambda q||^2 of one power step, and q = z / sqrt(||z||^2) for the
// next step in the same pass (the residual itself is not kept); lz points to
// the step's (lambda, ||z||^2) on the device.
struct MapPowerResid {
    static constexpr int NS = 1, NM = 0;
    double *partials;
    double *q;
    const double *z;
    const double *lz;
    __device__ void row(int i, double *s, double *) const
    {
        const double zi = z[i];
        const double r = zi - lz[0] * q[i];
        q[i] = zi / sqrt(lz[1]);
        s[0] += r * r;
    }
};

// workspace.c:26-43: diag(Q) and max |offdiag|
struct DiagInfo {
    static constexpr int NS = 0, NM = 1;
    double *partials;
    double *diag;
    const int *rp, *ci;
    const double *val;
    __device__ void row(int r, double *, double *mx) const
    {
        double d = 0.0, off = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            if (ci[k] == r) d += val[k];
            else off = fmax(off, fabs(val[k]));
        }
        diag[r] = d;
        mx[0] = fmax(mx[0], off);
    }
};

}  // namespace hprqp
