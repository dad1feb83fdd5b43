/*
 * CPU oracle for the HPR-QP iteration -- TEST INFRASTRUCTURE, never product.
 *
 * Restates the reference's CPU path with its floating-point operation order
 * kept, so that results can be compared bitwise with the reference built from
 * source (oracle/_ref). Every function cites the reference file:line it
 * follows. The cone (SOC) step has no CPU path in the reference
 * (src/shared/solver.c:250-253); it is restated here from the GPU semantics
 * in src/gpu/soc_common.cuh, soc_iteration.cuh and soc_residuals.cuh, with the
 * CPU's division-based scaling.
 *
 * Compile as ISO C11 (-std=c11 implies -ffp-contract=off): no fused
 * multiply-add, matching the reference CPU build (Makefile:107).
 */
#include "hprqp_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

/* ------------------------------------------------------------------------ */
/* CSR and vector helpers: src/shared/sparse.c:81-145, src/shared/vector.c */

typedef struct Csr {
    size_t nrow, ncol, nnz;
    size_t *ptr;
    size_t *ind;
    double *val;
} Csr;

static void csr_free(Csr *a)
{
    free(a->ptr);
    free(a->ind);
    free(a->val);
    memset(a, 0, sizeof(*a));
}

static int csr_alloc(Csr *a, size_t nrow, size_t ncol, size_t nnz)
{
    memset(a, 0, sizeof(*a));
    a->ptr = (size_t *)calloc(nrow + 1, sizeof(size_t));
    a->ind = (size_t *)calloc(nnz ? nnz : 1, sizeof(size_t));
    a->val = (double *)calloc(nnz ? nnz : 1, sizeof(double));
    if (!a->ptr || !a->ind || !a->val) {
        csr_free(a);
        return OR_ALLOC;
    }
    a->nrow = nrow;
    a->ncol = ncol;
    a->nnz = nnz;
    return OR_OK;
}

/* sparse.c:127-145: row sum, left to right, sum += v * x. */
void or_csr_matvec(size_t nrow, const size_t *ptr, const size_t *ind, const double *val,
                   const double *x, double *y)
{
    for (size_t r = 0; r < nrow; ++r) {
        double acc = 0.0;
        for (size_t k = ptr[r]; k < ptr[r + 1]; ++k) acc += val[k] * x[ind[k]];
        y[r] = acc;
    }
}

static void spmv(const Csr *a, const double *x, double *y)
{
    or_csr_matvec(a->nrow, a->ptr, a->ind, a->val, x, y);
}

/* sparse.c:81-125: count per column, prefix, scatter in row order (stable). */
void or_csr_transpose(size_t nrow, size_t ncol, const size_t *ptr, const size_t *ind, const double *val,
                      size_t *tptr, size_t *tind, double *tval)
{
    size_t nnz = ptr[nrow];
    size_t *fill = (size_t *)malloc((ncol ? ncol : 1) * sizeof(size_t));
    memset(tptr, 0, (ncol + 1) * sizeof(size_t));
    for (size_t k = 0; k < nnz; ++k) tptr[ind[k] + 1] += 1;
    for (size_t j = 0; j < ncol; ++j) tptr[j + 1] += tptr[j];
    memcpy(fill, tptr, ncol * sizeof(size_t));
    for (size_t r = 0; r < nrow; ++r) {
        for (size_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            size_t dst = fill[ind[k]]++;
            tind[dst] = r;
            tval[dst] = val[k];
        }
    }
    free(fill);
}

static int csr_transpose(Csr *t, const Csr *a)
{
    if (csr_alloc(t, a->ncol, a->nrow, a->nnz) != OR_OK) return OR_ALLOC;
    or_csr_transpose(a->nrow, a->ncol, a->ptr, a->ind, a->val, t->ptr, t->ind, t->val);
    return OR_OK;
}

/* vector.c:29-38: sequential dot. */
static double dot(const double *a, const double *b, size_t n)
{
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) acc += a[i] * b[i];
    return acc;
}

static double nrm2(const double *a, size_t n) { return sqrt(dot(a, a, n)); }

/* vector.c:43-56 */
static double nrm_inf(const double *a, size_t n)
{
    double best = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double v = fabs(a[i]);
        if (v > best) best = v;
    }
    return best;
}

static double clamp(double v, double lo, double hi)
{
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

static double wall_seconds(void)
{
    struct timeval tv;
    if (gettimeofday(&tv, NULL) != 0) return 0.0;
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

/* ------------------------------------------------------------------------ */
/* Model build and preprocess: src/shared/build.c:157-400                    */

typedef struct Model {
    size_t n, m;
    Csr Q, A, AT;
    double *c, *AL, *AU, *l, *u;
    double obj_constant;
    size_t nvc, nrc;
    size_t *var_off, *row_off;
    double *soc_rhs;
} Model;

static void model_free(Model *md)
{
    csr_free(&md->Q);
    csr_free(&md->A);
    csr_free(&md->AT);
    free(md->c);
    free(md->AL);
    free(md->AU);
    free(md->l);
    free(md->u);
    free(md->var_off);
    free(md->row_off);
    free(md->soc_rhs);
    memset(md, 0, sizeof(*md));
}

static double *dup_vec(const double *src, size_t n)
{
    double *d;
    if (n == 0) return NULL;
    d = (double *)malloc(n * sizeof(double));
    if (d && src) memcpy(d, src, n * sizeof(double));
    return d;
}

typedef struct Entry {
    size_t row, col;
    double val;
} Entry;

/* Stable ordering by (row, col): bucket by row (stable counting pass), then
 * a stable insertion/merge by column inside each row. Equal keys keep their
 * generation order, which is what the reference's qsort (glibc's merge sort)
 * yields for build.c:168 on these inputs. */
static void stable_sort_entries(Entry *e, size_t cnt, size_t nrow)
{
    size_t *start = (size_t *)calloc(nrow + 1, sizeof(size_t));
    Entry *tmp = (Entry *)malloc((cnt ? cnt : 1) * sizeof(Entry));
    for (size_t k = 0; k < cnt; ++k) start[e[k].row + 1] += 1;
    for (size_t r = 0; r < nrow; ++r) start[r + 1] += start[r];
    {
        size_t *fill = (size_t *)malloc((nrow ? nrow : 1) * sizeof(size_t));
        memcpy(fill, start, nrow * sizeof(size_t));
        for (size_t k = 0; k < cnt; ++k) tmp[fill[e[k].row]++] = e[k];
        free(fill);
    }
    for (size_t r = 0; r < nrow; ++r) {
        /* stable insertion sort by col within the row bucket */
        for (size_t i = start[r] + 1; i < start[r + 1]; ++i) {
            Entry key = tmp[i];
            size_t j = i;
            while (j > start[r] && tmp[j - 1].col > key.col) {
                tmp[j] = tmp[j - 1];
                --j;
            }
            tmp[j] = key;
        }
    }
    memcpy(e, tmp, cnt * sizeof(Entry));
    free(tmp);
    free(start);
}

/* build.c:157-219: drop |v| <= tol, merge equal (row,col), drop tiny sums. */
static int csr_from_entries(Csr *out, size_t nrow, size_t ncol, Entry *e, size_t cnt, double tol)
{
    size_t uniq = 0, kept = 0;
    stable_sort_entries(e, cnt, nrow);
    for (size_t k = 0; k < cnt; ++k) {
        if (fabs(e[k].val) <= tol) continue;
        if (uniq > 0 && e[k].row == e[uniq - 1].row && e[k].col == e[uniq - 1].col) {
            e[uniq - 1].val += e[k].val;
        } else {
            e[uniq++] = e[k];
        }
    }
    for (size_t k = 0; k < uniq; ++k)
        if (fabs(e[k].val) <= tol) e[k].val = 0.0;
    for (size_t k = 0; k < uniq; ++k)
        if (e[k].val != 0.0) e[kept++] = e[k];
    if (csr_alloc(out, nrow, ncol, kept) != OR_OK) return OR_ALLOC;
    for (size_t k = 0; k < kept; ++k) out->ptr[e[k].row + 1] += 1;
    for (size_t r = 0; r < nrow; ++r) out->ptr[r + 1] += out->ptr[r];
    for (size_t k = 0; k < kept; ++k) {
        out->ind[k] = e[k].col;
        out->val[k] = e[k].val;
    }
    return OR_OK;
}

/* build.c:221-252: Q <- 0.5 (Q + Q^T), entries generated (r,c),(c,r) per nnz. */
static int symmetrize(Csr *out, const Csr *q)
{
    size_t cnt = 0;
    int st;
    Entry *e = (Entry *)malloc((q->nnz > 0 ? 2 * q->nnz : 1) * sizeof(Entry));
    if (!e) return OR_ALLOC;
    for (size_t r = 0; r < q->nrow; ++r) {
        for (size_t k = q->ptr[r]; k < q->ptr[r + 1]; ++k) {
            double half = 0.5 * q->val[k];
            e[cnt].row = r;
            e[cnt].col = q->ind[k];
            e[cnt].val = half;
            ++cnt;
            e[cnt].row = q->ind[k];
            e[cnt].col = r;
            e[cnt].val = half;
            ++cnt;
        }
    }
    st = csr_from_entries(out, q->nrow, q->ncol, e, cnt, 1e-15);
    free(e);
    return st;
}

/* build.c:254-337: remove rows with no entries or with (-inf, +inf) bounds. */
static int drop_empty_free_rows(Model *md)
{
    size_t kept = 0, nnz = 0, pos = 0;
    size_t *keep;
    Csr na;
    double *nal, *nau;
    if (md->m == 0) return OR_OK;
    keep = (size_t *)malloc(md->m * sizeof(size_t));
    for (size_t r = 0; r < md->m; ++r) {
        size_t len = md->A.ptr[r + 1] - md->A.ptr[r];
        int unbounded = isinf(md->AL[r]) && md->AL[r] < 0.0 && isinf(md->AU[r]) && md->AU[r] > 0.0;
        if (len > 0 && !unbounded) {
            keep[kept++] = r;
            nnz += len;
        }
    }
    if (kept == md->m) {
        free(keep);
        return OR_OK;
    }
    if (csr_alloc(&na, kept, md->n, nnz) != OR_OK) {
        free(keep);
        return OR_ALLOC;
    }
    nal = kept ? (double *)malloc(kept * sizeof(double)) : NULL;
    nau = kept ? (double *)malloc(kept * sizeof(double)) : NULL;
    for (size_t r = 0; r < kept; ++r) {
        size_t old = keep[r];
        nal[r] = md->AL[old];
        nau[r] = md->AU[old];
        for (size_t k = md->A.ptr[old]; k < md->A.ptr[old + 1]; ++k) {
            na.ind[pos] = md->A.ind[k];
            na.val[pos] = md->A.val[k];
            ++pos;
        }
        na.ptr[r + 1] = pos;
    }
    csr_free(&md->A);
    csr_free(&md->AT);
    free(md->AL);
    free(md->AU);
    md->A = na;
    md->AL = nal;
    md->AU = nau;
    md->m = kept;
    free(keep);
    return csr_transpose(&md->AT, &md->A);
}

static int offsets_ok(size_t count, const size_t *off, size_t len)
{
    if (count == 0) return 1;
    if (count > len || off == NULL || off[0] >= len || off[count] != len) return 0;
    for (size_t k = 0; k < count; ++k)
        if (off[k] >= off[k + 1] || off[k + 1] > len) return 0;
    return 1;
}

/* build.c:51-77 */
static int validate_cones(const Model *md)
{
    if (!offsets_ok(md->nvc, md->var_off, md->n) || !offsets_ok(md->nrc, md->row_off, md->m)) return OR_INVALID;
    if (md->nvc) {
        for (size_t i = md->var_off[0]; i < md->n; ++i)
            if (md->l[i] != -INFINITY || md->u[i] != INFINITY) return OR_INVALID;
    }
    if (md->nrc) {
        if (!md->soc_rhs) return OR_INVALID;
        for (size_t i = 0; i < md->m; ++i) {
            if (!isfinite(md->soc_rhs[i])) return OR_INVALID;
            if (i < md->row_off[0]) {
                if (md->soc_rhs[i] != 0.0) return OR_INVALID;
            } else if (md->AL[i] != -INFINITY || md->AU[i] != INFINITY) {
                return OR_INVALID;
            }
        }
    }
    return OR_OK;
}

static int copy_csr(Csr *dst, size_t nrow, size_t ncol, size_t nnz, const size_t *ptr, const size_t *ind,
                    const double *val)
{
    if (ptr == NULL || ptr[0] != 0 || ptr[nrow] != nnz) return OR_INVALID;
    if (csr_alloc(dst, nrow, ncol, nnz) != OR_OK) return OR_ALLOC;
    memcpy(dst->ptr, ptr, (nrow + 1) * sizeof(size_t));
    if (nnz) {
        memcpy(dst->ind, ind, nnz * sizeof(size_t));
        memcpy(dst->val, val, nnz * sizeof(double));
    }
    return OR_OK;
}

/* build.c:339-400 and :504-592 */
static int model_build(Model *md, size_t n, size_t m,
                       size_t q_nnz, const size_t *q_ptr, const size_t *q_ind, const double *q_val,
                       size_t a_nnz, const size_t *a_ptr, const size_t *a_ind, const double *a_val,
                       const double *c, const double *AL, const double *AU,
                       const double *l, const double *u, double obj_constant,
                       size_t nvc, const size_t *var_off, size_t nrc, const size_t *row_off,
                       const double *soc_rhs)
{
    int st;
    Csr sym;
    memset(md, 0, sizeof(*md));
    if (c == NULL || l == NULL || u == NULL || (m > 0 && (AL == NULL || AU == NULL))) return OR_INVALID;
    md->n = n;
    md->m = m;
    if ((st = copy_csr(&md->Q, n, n, q_nnz, q_ptr, q_ind, q_val)) != OR_OK) return st;
    if ((st = copy_csr(&md->A, m, n, a_nnz, a_ptr, a_ind, a_val)) != OR_OK) return st;
    if ((st = csr_transpose(&md->AT, &md->A)) != OR_OK) return st;
    md->c = dup_vec(c, n);
    md->l = dup_vec(l, n);
    md->u = dup_vec(u, n);
    md->AL = dup_vec(AL, m);
    md->AU = dup_vec(AU, m);
    md->obj_constant = obj_constant;
    if (nvc || nrc) {
        if (!offsets_ok(nvc, var_off, n) || !offsets_ok(nrc, row_off, m) || (nrc && soc_rhs == NULL))
            return OR_INVALID;
        if (nvc) {
            md->var_off = (size_t *)malloc((nvc + 1) * sizeof(size_t));
            memcpy(md->var_off, var_off, (nvc + 1) * sizeof(size_t));
            md->nvc = nvc;
        }
        if (nrc) {
            md->row_off = (size_t *)malloc((nrc + 1) * sizeof(size_t));
            memcpy(md->row_off, row_off, (nrc + 1) * sizeof(size_t));
            md->nrc = nrc;
            md->soc_rhs = dup_vec(soc_rhs, m);
        }
        if ((st = validate_cones(md)) != OR_OK) return st;
    }
    /* preprocess_model: validation */
    for (size_t k = 0; k < md->Q.nnz; ++k)
        if (!isfinite(md->Q.val[k])) return OR_INVALID;
    for (size_t k = 0; k < md->A.nnz; ++k)
        if (!isfinite(md->A.val[k])) return OR_INVALID;
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(md->c[i]) || isnan(md->l[i]) || isnan(md->u[i]) || md->l[i] > md->u[i]) return OR_INVALID;
    for (size_t i = 0; i < m; ++i)
        if (isnan(md->AL[i]) || isnan(md->AU[i]) || md->AL[i] > md->AU[i]) return OR_INVALID;
    if ((st = validate_cones(md)) != OR_OK) return st;
    if ((st = symmetrize(&sym, &md->Q)) != OR_OK) return st;
    csr_free(&md->Q);
    md->Q = sym;
    if (md->nvc || md->nrc) return OR_OK;
    return drop_empty_free_rows(md);
}

/* ------------------------------------------------------------------------ */
/* Scaling: src/cpu/scale.c:10-318 (SOC block max: src/gpu/soc_common.cuh:57-64) */

typedef struct Scaling {
    double *row_norm, *col_norm;
    double b_scale, c_scale;
    double norm_b, norm_c, norm_b_org, norm_c_org;
} Scaling;

static double bound_mag(double al, double au)
{
    double a = isinf(al) ? 0.0 : fabs(al);
    double b = isinf(au) ? 0.0 : fabs(au);
    return a > b ? a : b;
}

static void cone_block_max(double *v, size_t count, const size_t *off)
{
    for (size_t k = 0; k < count; ++k) {
        double best = 0.0;
        for (size_t j = off[k]; j < off[k + 1]; ++j) best = fmax(best, v[j]);
        for (size_t j = off[k]; j < off[k + 1]; ++j) v[j] = best;
    }
}

/* scale.c:62-163: sqrt of column/row max-abs (sum_abs when use_sum). */
static void col_norms(double *out, const Model *md, int use_sum)
{
    for (size_t j = 0; j < md->n; ++j) out[j] = 0.0;
    for (size_t r = 0; r < md->A.nrow; ++r)
        for (size_t k = md->A.ptr[r]; k < md->A.ptr[r + 1]; ++k) {
            double v = fabs(md->A.val[k]);
            if (use_sum) out[md->A.ind[k]] += v;
            else if (v > out[md->A.ind[k]]) out[md->A.ind[k]] = v;
        }
    for (size_t r = 0; r < md->Q.nrow; ++r)
        for (size_t k = md->Q.ptr[r]; k < md->Q.ptr[r + 1]; ++k) {
            double v = fabs(md->Q.val[k]);
            if (use_sum) out[md->Q.ind[k]] += v;
            else if (v > out[md->Q.ind[k]]) out[md->Q.ind[k]] = v;
        }
    for (size_t j = 0; j < md->n; ++j) {
        out[j] = sqrt(out[j]);
        if (out[j] == 0.0) out[j] = 1.0;
    }
}

static void row_norms(double *out, const Model *md, int use_sum)
{
    for (size_t r = 0; r < md->A.nrow; ++r) {
        double acc = 0.0;
        for (size_t k = md->A.ptr[r]; k < md->A.ptr[r + 1]; ++k) {
            double v = fabs(md->A.val[k]);
            if (use_sum) acc += v;
            else if (v > acc) acc = v;
        }
        out[r] = sqrt(acc);
        if (out[r] == 0.0) out[r] = 1.0;
    }
}

/* scale.c:165-207 */
static void apply_scaling(Model *md, const double *rn, const double *cn)
{
    for (size_t r = 0; r < md->Q.nrow; ++r) {
        double d = cn[r];
        for (size_t k = md->Q.ptr[r]; k < md->Q.ptr[r + 1]; ++k) {
            md->Q.val[k] /= d;
            md->Q.val[k] /= cn[md->Q.ind[k]];
        }
    }
    for (size_t r = 0; r < md->A.nrow; ++r) {
        double d = rn[r];
        for (size_t k = md->A.ptr[r]; k < md->A.ptr[r + 1]; ++k) {
            md->A.val[k] /= d;
            md->A.val[k] /= cn[md->A.ind[k]];
        }
    }
    for (size_t j = 0; j < md->n; ++j) {
        md->c[j] /= cn[j];
        md->l[j] *= cn[j];
        md->u[j] *= cn[j];
    }
    for (size_t r = 0; r < md->m; ++r) {
        if (!isinf(md->AL[r])) md->AL[r] /= rn[r];
        if (!isinf(md->AU[r])) md->AU[r] /= rn[r];
        if (md->nrc) md->soc_rhs[r] /= rn[r];
    }
}

/* scale.c:231-318 */
static int scale_model(Model *md, int use_scaling, Scaling *sc)
{
    double *tr, *tc;
    memset(sc, 0, sizeof(*sc));
    sc->row_norm = (double *)malloc((md->m ? md->m : 1) * sizeof(double));
    sc->col_norm = (double *)malloc((md->n ? md->n : 1) * sizeof(double));
    tr = (double *)malloc((md->m ? md->m : 1) * sizeof(double));
    tc = (double *)malloc((md->n ? md->n : 1) * sizeof(double));
    for (size_t r = 0; r < md->m; ++r) sc->row_norm[r] = 1.0;
    for (size_t j = 0; j < md->n; ++j) sc->col_norm[j] = 1.0;
    sc->norm_b_org = 0.0;
    for (size_t r = 0; r < md->m; ++r) {
        double v = bound_mag(md->AL[r], md->AU[r]);
        if (v > sc->norm_b_org) sc->norm_b_org = v;
    }
    sc->norm_c_org = nrm_inf(md->c, md->n);
    if (use_scaling) {
        for (int pass = 0; pass < 11; ++pass) {
            int use_sum = pass == 10;
            col_norms(tc, md, use_sum);
            row_norms(tr, md, use_sum);
            if (md->nvc) cone_block_max(tc, md->nvc, md->var_off);
            if (md->nrc) cone_block_max(tr, md->nrc, md->row_off);
            for (size_t r = 0; r < md->m; ++r) sc->row_norm[r] *= tr[r];
            for (size_t j = 0; j < md->n; ++j) sc->col_norm[j] *= tc[j];
            apply_scaling(md, tr, tc);
        }
    }
    free(tr);
    free(tc);
    csr_free(&md->AT);
    if (csr_transpose(&md->AT, &md->A) != OR_OK) return OR_ALLOC;
    {
        double acc = 0.0;
        for (size_t r = 0; r < md->m; ++r) {
            double v = bound_mag(md->AL[r], md->AU[r]);
            acc += v * v;
        }
        sc->norm_b = sqrt(acc);
    }
    sc->norm_c = nrm2(md->c, md->n);
    sc->b_scale = 1.0;
    sc->c_scale = 1.0;
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Workspace: src/cpu/workspace.c:26-248                                     */

/* Equality-row (Gram) metric state; kind 0 when the metric is off. */
enum { GRAM_NONE = 0, GRAM_QP = 1, GRAM_ROW = 2 };
typedef struct Gram {
    int kind;
    size_t e;        /* metric rows */
    size_t *rows;    /* row index of each metric row */
    long *map;       /* row -> metric index, -1 for the other rows (m entries) */
    double *b;       /* right-hand side of each metric row */
    double *G;       /* the Gram matrix, dense row-major (kept for inspection) */
    double *L;       /* its Cholesky factor, dense row-major lower */
    Csr ET;          /* A_E^T: n rows, column = row index of A */
    double beta, backward, relative;
    size_t gnnz;     /* nonzeros of the assembled G */
    double *r, *step, *tmp;
} Gram;

typedef struct Work {
    size_t n, m;
    double sigma, lambda_max_A, lambda_max_Q;
    size_t power_steps_A, power_steps_Q; /* power-iteration products taken (0: not run) */
    int Q_is_diag, noC, to_check, soc;
    Gram g;
    double *x, *x_bar, *x_hat, *w, *w_bar, *y, *y_bar, *y_hat;
    double *Ax, *ATy, *ATy_bar, *ATdy, *Qw, *Qw_bar, *Qx, *dQw, *QATdy;
    double *dx, *dy, *dw, *Rp, *Rd, *s, *z_bar, *tempv;
    double *diag_Q, *fact1, *fact2, *fact, *fact_M;
    double *last_x, *last_y, *last_Qw, *last_w, *last_ATy;
} Work;

#define OR_WORK_N_VECTORS(X) X(x) X(x_bar) X(x_hat) X(w) X(w_bar) X(ATy) X(ATy_bar) X(ATdy) X(Qw) \
    X(Qw_bar) X(Qx) X(dQw) X(QATdy) X(dx) X(dw) X(Rd) X(z_bar) X(tempv) X(diag_Q) X(fact1) X(fact2) \
    X(fact) X(fact_M) X(last_x) X(last_Qw) X(last_w) X(last_ATy)
#define OR_WORK_M_VECTORS(X) X(y) X(y_bar) X(y_hat) X(Ax) X(dy) X(Rp) X(s) X(last_y)

static void gram_free(Gram *g)
{
    free(g->rows);
    free(g->map);
    free(g->b);
    free(g->G);
    free(g->L);
    csr_free(&g->ET);
    free(g->r);
    free(g->step);
    free(g->tmp);
    memset(g, 0, sizeof(*g));
}

static void work_free(Work *w)
{
    gram_free(&w->g);
#define OR_FREE(v) free(w->v);
    OR_WORK_N_VECTORS(OR_FREE)
    OR_WORK_M_VECTORS(OR_FREE)
#undef OR_FREE
    memset(w, 0, sizeof(*w));
}

static int work_alloc(Work *w, size_t n, size_t m)
{
    int bad = 0;
    w->n = n;
    w->m = m;
#define OR_ALLOC_N(v) w->v = (double *)calloc(n ? n : 1, sizeof(double)); bad |= w->v == NULL;
#define OR_ALLOC_M(v) w->v = (double *)calloc(m ? m : 1, sizeof(double)); bad |= w->v == NULL;
    OR_WORK_N_VECTORS(OR_ALLOC_N)
    OR_WORK_M_VECTORS(OR_ALLOC_M)
#undef OR_ALLOC_N
#undef OR_ALLOC_M
    return bad ? OR_ALLOC : OR_OK;
}

/* workspace.c:26-43 */
static int q_diag_check(const Csr *q, double *diag)
{
    int is_diag = 1;
    for (size_t r = 0; r < q->nrow; ++r) {
        diag[r] = 0.0;
        for (size_t k = q->ptr[r]; k < q->ptr[r + 1]; ++k) {
            if (q->ind[k] == r) diag[r] += q->val[k];
            else if (fabs(q->val[k]) > 1e-15) is_diag = 0;
        }
    }
    return is_diag;
}

/* workspace.c:78 and control.c:116-127 */
static void diag_factors(Work *w)
{
    for (size_t i = 0; i < w->n; ++i) {
        double inv = 1.0 / (1.0 + w->sigma * w->diag_Q[i]);
        w->fact2[i] = inv;
        w->fact[i] = w->sigma * inv;
        w->fact1[i] = w->sigma * w->diag_Q[i] * inv;
        w->fact_M[i] = w->sigma * w->sigma * inv;
    }
}

/* eigs.c:11-153: power iteration on A A^T (which=0) or Q (which=1); *steps
 * receives the number of products taken. */
static double power_iteration(const Model *md, int which, size_t *steps)
{
    size_t len = which ? md->n : md->m;
    double *z, *q, *tmp = NULL;
    double lam = 1.0;
    *steps = 0;
    if (len == 0) return 0.0;
    z = (double *)malloc(len * sizeof(double));
    q = (double *)malloc(len * sizeof(double));
    if (!which) tmp = (double *)malloc((md->n ? md->n : 1) * sizeof(double));
    for (size_t i = 0; i < len; ++i) z[i] = 1.0 + ((double)i + 1.0) / ((double)len + 1.0);
    for (size_t it = 0; it < 5000; ++it) {
        double qn = nrm2(z, len), zn, err;
        if (qn <= 1e-16) {
            lam = 0.0;
            break;
        }
        for (size_t i = 0; i < len; ++i) q[i] = z[i] / qn;
        if (which) {
            spmv(&md->Q, q, z);
        } else {
            spmv(&md->AT, q, tmp);
            spmv(&md->A, tmp, z);
        }
        *steps = it + 1;
        lam = dot(q, z, len);
        zn = nrm2(z, len);
        for (size_t i = 0; i < len; ++i) q[i] = z[i] - lam * q[i];
        err = nrm2(q, len) / (zn + fabs(lam) + 1e-16);
        if (err < 1e-4) break;
    }
    free(z);
    free(q);
    free(tmp);
    return lam < 0.0 ? 0.0 : lam;
}

static void gram_setup(Work *w, const Model *md, int mode);

/* workspace.c:191-248 (+ control.cu:189-220 for cone models) */
static int work_setup(Work *w, const Model *md, const OrParams *p)
{
    size_t free_cnt = 0;
    double nb = 0.0, nc, nb2 = 0.0;
    if (work_alloc(w, md->n, md->m) != OR_OK) return OR_ALLOC;
    w->soc = md->nvc || md->nrc;
    w->to_check = 1;
    for (size_t i = 0; i < md->n; ++i)
        if (isinf(md->l[i]) && isinf(md->u[i])) free_cnt += 1;
    w->noC = md->n > 0 && 5 * free_cnt > 4 * md->n;
    nc = nrm2(md->c, md->n);
    for (size_t r = 0; r < md->m; ++r) {
        double v = bound_mag(md->AL[r], md->AU[r]);
        nb2 += v * v;
    }
    nb = sqrt(nb2);
    if (md->nrc) nb = fmax(nb, nrm2(md->soc_rhs, md->m));
    w->sigma = (nb > 1e-16 && nc > 1e-16 && nb < 1e16 && nc < 1e16) ? nb / nc : 1.0;
    w->Q_is_diag = q_diag_check(&md->Q, w->diag_Q);
    if (w->Q_is_diag) diag_factors(w);
    /* eigs.c:155-185 */
    w->lambda_max_A = p->lambda_A_given > 0.0 ? p->lambda_A_given : power_iteration(md, 0, &w->power_steps_A) * 1.05;
    if (w->Q_is_diag) {
        w->lambda_max_Q = 0.0;
        for (size_t i = 0; i < md->n; ++i) {
            double d = w->soc ? fabs(w->diag_Q[i]) : w->diag_Q[i];
            if (d > w->lambda_max_Q) w->lambda_max_Q = d;
        }
    } else {
        w->lambda_max_Q = p->lambda_Q_given > 0.0 ? p->lambda_Q_given : power_iteration(md, 1, &w->power_steps_Q) * 1.05;
    }
    if (w->soc) {
        w->noC = 0;
        if (md->m > 0 && w->lambda_max_A <= 0.0) w->lambda_max_A = 1.0;
    }
    /* control.cu:222-239: the row metric may change sigma, so the diagonal
     * factors are formed again after it */
    gram_setup(w, md, p->gram_mode);
    if (w->g.kind && w->Q_is_diag) diag_factors(w);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Cone projection: src/gpu/soc_common.cuh:7-53                              */

void or_soc_project(double *out, const double *in, size_t len)
{
    double t = in[0], mx = 0.0, sum = 0.0, comp = 0.0, norm;
    for (size_t j = 1; j < len; ++j) mx = fmax(mx, fabs(in[j]));
    if (mx > 0.0) {
        for (size_t j = 1; j < len; ++j) {
            double sc = in[j] / mx;
            double term = sc * sc - comp;
            double nx = sum + term;
            comp = (nx - sum) - term;
            sum = nx;
        }
    }
    norm = mx * sqrt(sum);
    if (norm <= t) {
        for (size_t j = 0; j < len; ++j) out[j] = in[j];
    } else if (norm <= -t) {
        for (size_t j = 0; j < len; ++j) out[j] = 0.0;
    } else {
        double alpha = 0.5 + 0.5 * (t / norm);
        out[0] = 0.5 * norm + 0.5 * t;
        for (size_t j = 1; j < len; ++j) out[j] = alpha * in[j];
    }
}

/* interval prefix then cone blocks */
static void project_mixed(double *out, const double *in, const double *lo, const double *hi, size_t len,
                          size_t ncones, const size_t *off)
{
    size_t lin = ncones ? off[0] : len;
    for (size_t i = 0; i < lin; ++i) out[i] = fmin(fmax(in[i], lo[i]), hi[i]);
    for (size_t k = 0; k < ncones; ++k) or_soc_project(out + off[k], in + off[k], off[k + 1] - off[k]);
}

/* ------------------------------------------------------------------------ */
/* Equality-row (Gram) metric. The reference runs it on the GPU only:        */
/* src/gpu/qp_row_metric.cu:94-149 (plain QPs, gram_mode REQUEST),           */
/* src/gpu/row_metric.cu:68-130 (the homogeneous equality prefix of row-cone */
/* models) and src/shared/gram_policy.cpp:6-50 (the AUTO rule). With        */
/* A = [E; C], G = E E' and beta >= ||C||_2^2 majorised by row and column    */
/* sums, the y update of the metric rows becomes y - G^{-1}(E t - b)/(2 sigma) */
/* and the other rows keep theirs with lambda_A = 2 beta. The reference      */
/* factors G with a sparse direct solver; the oracle uses a dense Cholesky.  */

/* gram_policy.cpp:6-50 on the built, unscaled model. */
static int gram_auto_eligible(const Model *md)
{
    size_t e, n = md->n, g;
    size_t *degree, *location, *state;
    double *q;
    int ok = 1;
    if (md->nvc || !md->nrc) return 0;
    e = md->row_off[0];
    g = (size_t)sqrt((double)e);
    if (g < 2 || g * g != e || e > n) return 0;
    for (size_t k = 0; k < md->nrc; ++k)
        if (md->row_off[k + 1] - md->row_off[k] != 3) return 0;
    for (size_t i = 0; i < e; ++i)
        if (md->AL[i] != 0.0 || md->AU[i] != 0.0 || md->soc_rhs[i] != 0.0) return 0;
    degree = (size_t *)calloc(n ? n : 1, sizeof(size_t));
    location = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
    state = (size_t *)malloc((e ? e : 1) * sizeof(size_t));
    q = (double *)calloc(n ? n : 1, sizeof(double));
    if (!degree || !location || !state || !q) ok = 0;
    if (ok) {
        size_t multi = 0;
        for (size_t j = 0; j < n; ++j) location[j] = e;
        for (size_t i = 0; i < e; ++i) state[i] = n;
        for (size_t k = 0; k < md->A.ptr[e]; ++k) degree[md->A.ind[k]] += 1;
        for (size_t j = 0; j < n; ++j) multi += degree[j] > 1;
        if (multi != e) ok = 0;
        /* each stencil row owns one shared column, with coefficient 4 */
        for (size_t i = 0; ok && i < e; ++i) {
            for (size_t k = md->A.ptr[i]; ok && k < md->A.ptr[i + 1]; ++k) {
                size_t j = md->A.ind[k];
                if (degree[j] > 1 && md->A.val[k] > 0.0) {
                    if (state[i] != n || location[j] != e || md->A.val[k] != 4.0) ok = 0;
                    state[i] = j;
                    location[j] = i;
                }
            }
            if (state[i] == n) ok = 0;
        }
        /* the shared columns form the five-point stencil of a g x g grid */
        for (size_t i = 0; ok && i < e; ++i) {
            size_t count = 0;
            for (size_t k = md->A.ptr[i]; ok && k < md->A.ptr[i + 1]; ++k) {
                size_t j = location[md->A.ind[k]];
                int nb;
                if (j == e) continue;
                nb = (i >= g && j == i - g) || (i + g < e && j == i + g) || (i % g && j == i - 1) ||
                     (i % g + 1 < g && j == i + 1);
                if (md->A.val[k] != (i == j ? 4.0 : (nb ? -1.0 : 0.0))) ok = 0;
                count += 1;
            }
            if (ok && count != (size_t)(1 + (i >= g) + (i + g < e) + (i % g > 0) + (i % g + 1 < g))) ok = 0;
        }
        /* Q diagonal with a positive finite entry on every state column */
        for (size_t i = 0; ok && i < n; ++i)
            for (size_t k = md->Q.ptr[i]; ok && k < md->Q.ptr[i + 1]; ++k) {
                if (md->Q.ind[k] != i && md->Q.val[k] != 0.0) ok = 0;
                if (md->Q.ind[k] == i) q[i] += md->Q.val[k];
            }
        for (size_t i = 0; ok && i < e; ++i)
            if (!(q[state[i]] > 0.0) || !isfinite(q[state[i]])) ok = 0;
    }
    free(degree);
    free(location);
    free(state);
    free(q);
    return ok;
}

/* Dense Cholesky G = L L' (row-major, lower); 0 when a pivot is not
 * positive and finite. */
static int gram_cholesky(const double *G, double *L, size_t e)
{
    memset(L, 0, e * e * sizeof(double));
    for (size_t j = 0; j < e; ++j) {
        double d = G[j * e + j];
        for (size_t t = 0; t < j; ++t) d -= L[j * e + t] * L[j * e + t];
        if (!(d > 0.0) || !isfinite(d)) return 0;
        d = sqrt(d);
        L[j * e + j] = d;
        for (size_t i = j + 1; i < e; ++i) {
            double v = G[i * e + j];
            for (size_t t = 0; t < j; ++t) v -= L[i * e + t] * L[j * e + t];
            L[i * e + j] = v / d;
        }
    }
    return 1;
}

/* x = G^{-1} r by forward then backward substitution with L. */
static void gram_solve(const Gram *g, const double *r, double *x)
{
    size_t e = g->e;
    for (size_t i = 0; i < e; ++i) {
        double v = r[i];
        for (size_t t = 0; t < i; ++t) v -= g->L[i * e + t] * x[t];
        x[i] = v / g->L[i * e + i];
    }
    for (size_t i = e; i-- > 0;) {
        double v = x[i];
        for (size_t t = i + 1; t < e; ++t) v -= g->L[t * e + i] * x[t];
        x[i] = v / g->L[i * e + i];
    }
}

typedef struct GramPair {
    size_t col, seq;
    double val;
} GramPair;

static int gram_pair_cmp(const void *a, const void *b)
{
    const GramPair *p = (const GramPair *)a, *q = (const GramPair *)b;
    if (p->col != q->col) return p->col < q->col ? -1 : 1;
    return p->seq < q->seq ? -1 : (p->seq > q->seq);
}

/* Shared by both metrics: given the metric rows (g->rows, g->map, g->e) and
 * beta, build E' and G (qp_row_metric.cu:109-116 / row_metric.cu:83-93), the
 * factor and its validation (:120-127 / :98-103). Returns 1 when accepted. */
static int gram_build(Gram *g, const Model *md)
{
    size_t e = g->e, n = md->n, nnz = 0;
    size_t *cnt;
    unsigned long long work = 0;
    GramPair *pairs = NULL;
    size_t cap = 0;
    double *t = NULL, *x = NULL, err = 0.0, an = 0.0, xn = 0.0, bn = 0.0;
    cnt = (size_t *)calloc(n + 1, sizeof(size_t));
    if (!cnt) return 0;
    for (size_t q = 0; q < e; ++q) {
        size_t i = g->rows[q];
        for (size_t k = md->A.ptr[i]; k < md->A.ptr[i + 1]; ++k) cnt[md->A.ind[k] + 1] += 1;
    }
    for (size_t j = 0; j < n; ++j) work += (unsigned long long)cnt[j + 1] * cnt[j + 1];
    if (work > 100000000ULL) {
        free(cnt);
        return 0;
    }
    for (size_t j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
    if (csr_alloc(&g->ET, n, md->m, cnt[n]) != OR_OK) {
        free(cnt);
        return 0;
    }
    memcpy(g->ET.ptr, cnt, (n + 1) * sizeof(size_t));
    for (size_t q = 0; q < e; ++q) {
        size_t i = g->rows[q];
        for (size_t k = md->A.ptr[i]; k < md->A.ptr[i + 1]; ++k) {
            size_t z = cnt[md->A.ind[k]]++;
            g->ET.ind[z] = i;
            g->ET.val[z] = md->A.val[k];
        }
    }
    free(cnt);
    g->G = (double *)calloc(e * e, sizeof(double));
    g->L = (double *)malloc(e * e * sizeof(double));
    g->r = (double *)calloc(e, sizeof(double));
    g->step = (double *)calloc(e, sizeof(double));
    g->tmp = (double *)calloc(n ? n : 1, sizeof(double));
    if (!g->G || !g->L || !g->r || !g->step || !g->tmp) return 0;
    /* row q of G: products of row i with every metric row sharing a column,
     * gathered in entry order and summed per column in that order */
    for (size_t q = 0; q < e; ++q) {
        size_t i = g->rows[q], cntp = 0;
        for (size_t k = md->A.ptr[i]; k < md->A.ptr[i + 1]; ++k) {
            size_t j = md->A.ind[k];
            size_t need = cntp + (g->ET.ptr[j + 1] - g->ET.ptr[j]);
            if (need > cap) {
                GramPair *np;
                cap = need * 2;
                np = (GramPair *)realloc(pairs, cap * sizeof(GramPair));
                if (!np) {
                    free(pairs);
                    return 0;
                }
                pairs = np;
            }
            for (size_t z = g->ET.ptr[j]; z < g->ET.ptr[j + 1]; ++z) {
                pairs[cntp].col = (size_t)g->map[g->ET.ind[z]];
                pairs[cntp].seq = cntp;
                pairs[cntp].val = md->A.val[k] * g->ET.val[z];
                cntp += 1;
            }
        }
        if (cntp) qsort(pairs, cntp, sizeof(GramPair), gram_pair_cmp);
        for (size_t z = 0; z < cntp;) {
            size_t col = pairs[z].col;
            double sum = 0.0;
            do {
                sum += pairs[z++].val;
            } while (z < cntp && pairs[z].col == col);
            if (sum != 0.0) {
                g->G[q * e + col] = sum;
                nnz += 1;
            }
        }
    }
    free(pairs);
    g->gnnz = nnz;
    if (!gram_cholesky(g->G, g->L, e)) return 0;
    /* validation solve (qp_row_metric.cu:120-127) */
    t = (double *)calloc(e, sizeof(double));
    x = (double *)calloc(e, sizeof(double));
    if (!t || !x) {
        free(t);
        free(x);
        return 0;
    }
    for (size_t i = 0; i < e; ++i) t[i] = sin(.173 * (double)(i + 1)) + .2 * cos(.719 * (double)(i + 1));
    gram_solve(g, t, x);
    for (size_t i = 0; i < e; ++i) {
        long double z = 0.0L;
        double rs = 0.0;
        if (!isfinite(x[i])) {
            free(t);
            free(x);
            return 0;
        }
        for (size_t k = 0; k < e; ++k) {
            double v = g->G[i * e + k];
            if (v != 0.0) {
                z += (long double)v * x[k];
                rs += fabs(v);
            }
        }
        err = fmax(err, (double)fabsl(z - (long double)t[i]));
        an = fmax(an, rs);
        xn = fmax(xn, fabs(x[i]));
        bn = fmax(bn, fabs(t[i]));
    }
    free(t);
    free(x);
    g->backward = err / (an * xn + bn);
    g->relative = err / bn;
    return g->backward <= 1e-10 && (e > 5000 || g->relative <= 1e-8);
}

/* qp_row_metric.cu:94-149: plain QPs, every row with AL == AU finite. */
static int gram_setup_qp(Work *w, const Model *md)
{
    Gram *g = &w->g;
    size_t m = md->m, n = md->n;
    double *cols, rowmax = 0.0, cmax = 0.0, normb, normc, *sol;
    long double rest = 0.0L, bq = 0.0L;
    if (!m || !n || md->A.nnz == 0) return 0;
    g->map = (long *)malloc(m * sizeof(long));
    g->rows = (size_t *)malloc(m * sizeof(size_t));
    g->b = (double *)malloc(m * sizeof(double));
    cols = (double *)calloc(n, sizeof(double));
    if (!g->map || !g->rows || !g->b || !cols) {
        free(cols);
        return 0;
    }
    for (size_t i = 0; i < m; ++i) {
        g->map[i] = -1;
        if (isfinite(md->AL[i]) && md->AL[i] == md->AU[i]) {
            g->map[i] = (long)g->e;
            g->rows[g->e] = i;
            g->b[g->e] = md->AL[i];
            g->e += 1;
        }
    }
    if (g->e == 0 || g->e > n) {
        free(cols);
        return 0;
    }
    for (size_t i = 0; i < m; ++i) {
        double row = 0.0, lo, hi, v;
        if (g->map[i] >= 0) continue;
        for (size_t k = md->A.ptr[i]; k < md->A.ptr[i + 1]; ++k) {
            row += fabs(md->A.val[k]);
            cols[md->A.ind[k]] += fabs(md->A.val[k]);
        }
        rowmax = fmax(rowmax, row);
        lo = isfinite(md->AL[i]) ? fabs(md->AL[i]) : 0.0;
        hi = isfinite(md->AU[i]) ? fabs(md->AU[i]) : 0.0;
        v = fmax(lo, hi);
        rest += (long double)v * v;
    }
    for (size_t j = 0; j < n; ++j) cmax = fmax(cmax, cols[j]);
    free(cols);
    g->beta = (1 + 1e-12) * rowmax * cmax;
    if (g->beta == 0.0) g->beta = 1.0;
    if (!isfinite(g->beta)) return 0;
    if (!gram_build(g, md)) return 0;
    /* sigma from the metric norm of b (:128-133) */
    sol = (double *)malloc(g->e * sizeof(double));
    if (!sol) return 0;
    gram_solve(g, g->b, sol);
    for (size_t q = 0; q < g->e; ++q) bq += (long double)g->b[q] * sol[q];
    free(sol);
    if (!isfinite((double)bq) || bq < -1e-10L) return 0;
    normb = sqrt((double)((bq > 0.0L ? bq : 0.0L) + rest / g->beta));
    normc = nrm2(md->c, n);
    w->sigma = (normb > 1e-16 && normc > 1e-16 && normb < 1e16 && normc < 1e16) ? normb / normc : 1.0;
    w->lambda_max_A = 2.0 * g->beta;
    w->noC = 0;
    g->kind = GRAM_QP;
    return 1;
}

/* row_metric.cu:68-130: the linear prefix of a row-cone model, all
 * homogeneous equalities; diagonal Q, no variable cones. */
static int gram_setup_row(Work *w, const Model *md)
{
    Gram *g = &w->g;
    size_t m = md->m, n = md->n, l;
    double *cols, rowmax = 0.0, cmax = 0.0;
    if (!md->nrc || md->nvc || !w->Q_is_diag || w->noC) return 0;
    l = md->row_off[0];
    if (!l) return 0;
    for (size_t i = 0; i < l; ++i)
        if (md->AL[i] != 0.0 || md->AU[i] != 0.0 || md->soc_rhs[i] != 0.0) return 0;
    cols = (double *)calloc(n ? n : 1, sizeof(double));
    if (!cols) return 0;
    for (size_t i = l; i < m; ++i) {
        double row = 0.0;
        for (size_t k = md->A.ptr[i]; k < md->A.ptr[i + 1]; ++k) {
            row += fabs(md->A.val[k]);
            cols[md->A.ind[k]] += fabs(md->A.val[k]);
        }
        rowmax = fmax(rowmax, row);
    }
    for (size_t j = 0; j < n; ++j) cmax = fmax(cmax, cols[j]);
    free(cols);
    g->beta = (1 + 1e-12) * rowmax * cmax;
    if (!(g->beta > 0.0)) return 0;
    g->e = l;
    g->map = (long *)malloc(m * sizeof(long));
    g->rows = (size_t *)malloc(l * sizeof(size_t));
    g->b = (double *)calloc(l, sizeof(double));
    if (!g->map || !g->rows || !g->b) return 0;
    for (size_t i = 0; i < m; ++i) g->map[i] = i < l ? (long)i : -1;
    for (size_t i = 0; i < l; ++i) g->rows[i] = i;
    if (!gram_build(g, md)) return 0;
    w->sigma /= sqrt(g->beta);
    w->lambda_max_A = 2.0 * g->beta;
    g->kind = GRAM_ROW;
    return 1;
}

/* control.cu:222-225 (setup) with the declined metric leaving the workspace
 * as it was (qp_row_metric.cu:140-147, row_metric.cu:133-146). */
static void gram_setup(Work *w, const Model *md, int mode)
{
    double sigma = w->sigma, lam = w->lambda_max_A;
    int noC = w->noC, ok = 0;
    if (mode == 0) return;
    if (!md->nvc && !md->nrc) ok = gram_setup_qp(w, md);
    else if (md->nrc) ok = gram_setup_row(w, md);
    if (!ok) {
        gram_free(&w->g);
        w->sigma = sigma;
        w->lambda_max_A = lam;
        w->noC = noC;
    }
}

/* The metric's y_bar on its rows (qp_row_metric.cu:55-63, row_metric.cu:46-55). */
static void gram_step(Work *w)
{
    Gram *g = &w->g;
    for (size_t q = 0; q < g->e; ++q) g->r[q] = w->Ax[g->rows[q]] - g->b[q];
    gram_solve(g, g->r, g->step);
}

/* 2 ||E' dy_E||^2 + 2 beta ||dy_rest||^2 (qp_row_metric.cu:64-73,
 * row_metric.cu:56-66). */
static double gram_quadratic(Work *w)
{
    Gram *g = &w->g;
    double eq, rest = 0.0;
    spmv(&g->ET, w->dy, g->tmp);
    eq = dot(g->tmp, g->tmp, w->n);
    for (size_t i = 0; i < w->m; ++i)
        if (g->map[i] < 0) rest += w->dy[i] * w->dy[i];
    return 2.0 * eq + 2.0 * g->beta * rest;
}

/* ------------------------------------------------------------------------ */
/* Iteration: src/cpu/iteration.c:19-293                                     */

/* iteration.c:19-83 */
static void step_zxw1_noC(Work *w, const Model *md, double h1, double h2)
{
    double f2s = 1.0 / (1.0 + w->sigma * w->lambda_max_Q);
    double f1s = 1.0 - f2s;
    if (w->m > 0) spmv(&md->AT, w->y, w->ATy);
    else for (size_t i = 0; i < w->n; ++i) w->ATy[i] = 0.0;
    for (size_t i = 0; i < w->n; ++i) {
        double tmp = -w->Qw[i] + w->ATy[i] - md->c[i];
        double zr = w->x[i] + w->sigma * tmp;
        double xb = clamp(zr, md->l[i], md->u[i]);
        double xh = 2.0 * xb - w->x[i];
        double xn = h2 * xh + h1 * w->last_x[i];
        double wb = w->Q_is_diag ? w->fact1[i] * w->w[i] + w->fact2[i] * xh : f1s * w->w[i] + f2s * xh;
        double wh = 2.0 * wb - w->w[i];
        double wn = h2 * wh + h1 * w->last_w[i];
        if (w->to_check) {
            w->dx[i] = xb - w->x[i];
            w->dw[i] = wb - w->w[i];
            w->z_bar[i] = (xb - zr) / w->sigma;
        }
        w->x_bar[i] = xb;
        w->x_hat[i] = xh;
        w->x[i] = xn;
        w->w_bar[i] = wb;
        w->w[i] = wn;
    }
    spmv(&md->Q, w->w_bar, w->Qw_bar);
    for (size_t i = 0; i < w->n; ++i) {
        double qw = w->Qw[i], qwb = w->Qw_bar[i];
        double qwh = 2.0 * qwb - qw;
        w->Qw[i] = h2 * qwh + h1 * w->last_Qw[i];
        w->tempv[i] = w->x_hat[i] + w->sigma * (qw - qwb);
    }
}

/* iteration.c:85-123 */
static void step_zxw1(Work *w, const Model *md, double h1, double h2)
{
    double f2s = 1.0 / (1.0 + w->sigma * w->lambda_max_Q);
    double f1s = 1.0 - f2s;
    spmv(&md->Q, w->w, w->Qw);
    for (size_t i = 0; i < w->n; ++i) {
        double tmp = -w->Qw[i] + w->ATy[i] - md->c[i];
        double zr = w->x[i] + w->sigma * tmp;
        double xb = clamp(zr, md->l[i], md->u[i]);
        double xh = 2.0 * xb - w->x[i];
        double xn = h2 * xh + h1 * w->last_x[i];
        double wb = w->Q_is_diag ? w->fact1[i] * w->w[i] + w->fact2[i] * xh : f1s * w->w[i] + f2s * xh;
        if (w->to_check) {
            w->dx[i] = xb - w->x[i];
            w->z_bar[i] = (xb - zr) / w->sigma;
        }
        w->x_bar[i] = xb;
        w->x_hat[i] = xh;
        w->x[i] = xn;
        w->w_bar[i] = wb;
    }
}

/* iteration.c:125-216 (y update; noC skips the Q w_bar / tempv part) */
static int step_y(Work *w, const Model *md, double h1, double h2, int noC)
{
    double f1, f2;
    if (w->m == 0) return OR_OK;
    if (!noC) {
        spmv(&md->Q, w->w_bar, w->Qw_bar);
        for (size_t i = 0; i < w->n; ++i) w->tempv[i] = w->x_hat[i] + w->sigma * (w->Qw[i] - w->Qw_bar[i]);
    }
    spmv(&md->A, w->tempv, w->Ax);
    f1 = w->lambda_max_A * w->sigma;
    if (f1 <= 0.0) return OR_INVALID;
    f2 = 1.0 / f1;
    for (size_t i = 0; i < w->m; ++i) {
        double sr = w->Ax[i] - f1 * w->y[i];
        double sp = clamp(sr, md->AL[i], md->AU[i]);
        double yb = f2 * (sp - sr);
        double yh = 2.0 * yb - w->y[i];
        double yn = h2 * yh + h1 * w->last_y[i];
        if (w->to_check) {
            w->s[i] = sp;
            w->dy[i] = yb - w->y[i];
        }
        w->y_bar[i] = yb;
        w->y_hat[i] = yh;
        w->y[i] = yn;
    }
    return OR_OK;
}

/* iteration.c:218-248 (and the full-write GPU variant, iteration.cu:573-603) */
static void step_w2(Work *w, const Model *md, double h1, double h2, int full)
{
    double fs = w->sigma / (1.0 + w->sigma * w->lambda_max_Q);
    if (w->m > 0) spmv(&md->AT, w->y_bar, w->ATy_bar);
    for (size_t i = 0; i < w->n; ++i) {
        double f = w->Q_is_diag ? w->fact[i] : fs;
        double wbn = w->w_bar[i] + f * (w->ATy_bar[i] - w->ATy[i]);
        double wn = h1 * w->last_w[i] + h2 * (2.0 * wbn - w->w[i]);
        double atyn = h1 * w->last_ATy[i] + h2 * (2.0 * w->ATy_bar[i] - w->ATy[i]);
        if (w->to_check || full) {
            w->dw[i] = wbn - w->w[i];
            w->ATdy[i] = w->ATy_bar[i] - w->ATy[i];
        }
        w->w_bar[i] = wbn;
        w->w[i] = wn;
        w->ATy[i] = atyn;
    }
}

/* soc_iteration.cuh:45-82: the generic cone step with full state writes. */
static void step_soc(Work *w, const Model *md, double a, double b)
{
    size_t n = w->n, m = w->m;
    spmv(&md->Q, w->w, w->Qw);
    for (size_t i = 0; i < n; ++i) w->tempv[i] = w->x[i] + w->sigma * (-w->Qw[i] + w->ATy[i] - md->c[i]);
    project_mixed(w->x_bar, w->tempv, md->l, md->u, n, md->nvc, md->var_off);
    for (size_t i = 0; i < n; ++i) {
        double x = w->x[i], xb = w->x_bar[i], xh = 2.0 * xb - x;
        double f2 = w->Q_is_diag ? w->fact2[i] : 1.0 / (1.0 + w->sigma * w->lambda_max_Q);
        double f1 = w->Q_is_diag ? w->fact1[i] : 1.0 - f2;
        w->dx[i] = xb - x;
        w->z_bar[i] = (xb - w->tempv[i]) / w->sigma;
        w->x_hat[i] = xh;
        w->x[i] = a * w->last_x[i] + b * xh;
        w->w_bar[i] = f1 * w->w[i] + f2 * xh;
    }
    spmv(&md->Q, w->w_bar, w->Qw_bar);
    for (size_t i = 0; i < n; ++i) w->tempv[i] = w->x_hat[i] + w->sigma * (w->Qw[i] - w->Qw_bar[i]);
    if (m > 0) {
        double tau = w->sigma * w->lambda_max_A;
        spmv(&md->A, w->tempv, w->Ax);
        for (size_t i = 0; i < m; ++i) w->Rp[i] = w->Ax[i] - (md->nrc ? md->soc_rhs[i] : 0.0) - tau * w->y[i];
        project_mixed(w->s, w->Rp, md->AL, md->AU, m, md->nrc, md->row_off);
        if (w->g.kind) gram_step(w);
        for (size_t i = 0; i < m; ++i) {
            double y = w->y[i];
            double yb = (w->g.kind && w->g.map[i] >= 0) ? y - w->g.step[w->g.map[i]] / (2.0 * w->sigma)
                                                         : (w->s[i] - w->Rp[i]) / tau;
            w->dy[i] = yb - y;
            w->y_bar[i] = yb;
            w->y[i] = a * w->last_y[i] + b * (2.0 * yb - y);
        }
    }
    step_w2(w, md, a, b, 1);
}

/* iteration.c:250-293 */
static int iteration_step(Work *w, const Model *md, size_t inner)
{
    double h1 = 1.0 / ((double)inner + 2.0);
    double h2 = 1.0 - h1;
    int st;
    if (w->soc || w->g.kind) {
        step_soc(w, md, h1, h2);
        return OR_OK;
    }
    if (w->noC) {
        step_zxw1_noC(w, md, h1, h2);
        return step_y(w, md, h1, h2, 1);
    }
    step_zxw1(w, md, h1, h2);
    if ((st = step_y(w, md, h1, h2, 0)) != OR_OK) return st;
    step_w2(w, md, h1, h2, 0);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Residuals: src/cpu/residuals.c:46-189 (cones: soc_residuals.cuh:27-78)    */

typedef struct Resid {
    int updated;
    double err_rp, err_rd, kkt, pobj, gap, dobj;
} Resid;

static void residuals(Work *w, const Model *md, const Scaling *sc, Resid *r, size_t iter)
{
    size_t n = w->n, m = w->m;
    double mqx = 0.0, maty = 0.0, max_ax = 0.0, scale = sc->b_scale * sc->c_scale;
    double cx, xqx, zx, ys = 0.0, yb = 0.0;
    memset(r, 0, sizeof(*r));
    spmv(&md->Q, w->x_bar, w->Qx);
    if (m > 0) {
        spmv(&md->A, w->x_bar, w->Ax);
        spmv(&md->AT, w->y_bar, w->ATdy);
    } else {
        for (size_t i = 0; i < n; ++i) w->ATdy[i] = 0.0;
    }
    cx = dot(md->c, w->x_bar, n);
    xqx = dot(w->x_bar, w->Qx, n);
    zx = dot(w->z_bar, w->x_bar, n);
    if (w->g.kind == GRAM_QP)
        for (size_t i = 0; i < m; ++i) w->s[i] = clamp(w->Ax[i], md->AL[i], md->AU[i]);
    if (m > 0) ys = dot(w->y_bar, w->s, m);
    if (md->nrc) yb = dot(w->y_bar, md->soc_rhs, m);
    r->pobj = scale * (cx + 0.5 * xqx) + md->obj_constant;
    r->dobj = w->soc ? scale * (-0.5 * xqx + zx + ys + yb) + md->obj_constant
                     : scale * (-0.5 * xqx + zx + ys) + md->obj_constant;
    r->gap = fabs(r->pobj - r->dobj) / (1.0 + fmax(fabs(r->pobj), fabs(r->dobj)));
    for (size_t i = 0; i < n; ++i) {
        double f = sc->col_norm[i] * sc->c_scale;
        double sq = w->Qx[i] * f, sa = w->ATdy[i] * f;
        w->Rd[i] = (w->Qx[i] + md->c[i] - w->ATdy[i] - w->z_bar[i]) * f;
        if (fabs(sq) > mqx) mqx = fabs(sq);
        if (fabs(sa) > maty) maty = fabs(sa);
    }
    r->err_rd = nrm_inf(w->Rd, n) / (1.0 + fmax(fmax(sc->norm_c_org, maty), mqx));
    if (w->soc) {
        /* soc_residuals.cuh:27-78: the feasibility projection is taken on
         * Ax - rhs; the variable-domain violation is audited every check. */
        double *arg = (double *)malloc((m > n ? m : n) * sizeof(double) + 8);
        double *proj = (double *)malloc((m > n ? m : n) * sizeof(double) + 8);
        double viol = 0.0;
        if (m > 0) {
            for (size_t i = 0; i < m; ++i) arg[i] = w->Ax[i] - (md->nrc ? md->soc_rhs[i] : 0.0);
            project_mixed(proj, arg, md->AL, md->AU, m, md->nrc, md->row_off);
            for (size_t i = 0; i < m; ++i) {
                double f = sc->row_norm[i] * sc->b_scale;
                w->Rp[i] = (proj[i] - arg[i]) * f;
                if (fabs(w->Ax[i] * f) > max_ax) max_ax = fabs(w->Ax[i] * f);
            }
            r->err_rp = nrm_inf(w->Rp, m) / (1.0 + fmax(sc->norm_b_org, max_ax));
        }
        project_mixed(proj, w->x_bar, md->l, md->u, n, md->nvc, md->var_off);
        for (size_t i = 0; i < n; ++i) {
            double v = fabs((proj[i] - w->x_bar[i]) * sc->b_scale / sc->col_norm[i]);
            if (v > viol) viol = v;
        }
        r->err_rp = fmax(r->err_rp, viol);
        free(arg);
        free(proj);
    } else {
        if (m > 0) {
            for (size_t i = 0; i < m; ++i) {
                double p = clamp(w->Ax[i], md->AL[i], md->AU[i]);
                double f = sc->row_norm[i] * sc->b_scale;
                w->Rp[i] = (p - w->Ax[i]) * f;
                if (fabs(w->Ax[i] * f) > max_ax) max_ax = fabs(w->Ax[i] * f);
            }
            r->err_rp = nrm_inf(w->Rp, m) / (1.0 + fmax(sc->norm_b_org, max_ax));
        }
        if (iter == 0) {
            /* residuals.c:58-78 */
            double worst = 0.0;
            for (size_t i = 0; i < n; ++i) {
                double lo = md->l[i] - w->x_bar[i], hi = w->x_bar[i] - md->u[i], v = 0.0;
                if (lo > 0.0) v += lo;
                if (hi > 0.0) v += hi;
                w->dx[i] = v * (sc->b_scale / sc->col_norm[i]);
                if (w->dx[i] > worst) worst = w->dx[i];
            }
            if (worst > r->err_rp) r->err_rp = worst;
        }
    }
    r->kkt = fmax(fmax(r->err_rp, r->err_rd), r->gap);
    r->updated = 1;
}

/* ------------------------------------------------------------------------ */
/* Control: src/cpu/control.c:20-457                                         */

typedef struct Restart {
    int flag, first;
    double last_gap, current_gap, save_gap;
    size_t inner, sufficient, necessary, long_restart, times;
    double weighted_norm, best_gap, best_sigma;
} Restart;

static void restart_reset(Restart *r)
{
    memset(r, 0, sizeof(*r));
    r->first = 1;
    r->save_gap = INFINITY;
    r->current_gap = INFINITY;
    r->last_gap = INFINITY;
    r->best_gap = INFINITY;
    r->best_sigma = 1.0;
    r->weighted_norm = INFINITY;
}

/* control.c:155-229 */
static double m_norm(Work *w, const Model *md)
{
    size_t n = w->n, m = w->m;
    double m1 = 0.0, m2, m3 = 0.0;
    m2 = (1.0 / w->sigma) * dot(w->dx, w->dx, n);
    spmv(&md->Q, w->dw, w->dQw);
    m2 -= 2.0 * dot(w->dQw, w->dx, n);
    if (m > 0) {
        m1 = w->g.kind ? w->sigma * gram_quadratic(w) : w->sigma * w->lambda_max_A * dot(w->dy, w->dy, m);
        spmv(&md->AT, w->dy, w->ATdy);
        spmv(&md->Q, w->ATdy, w->QATdy);
        m1 -= 2.0 * w->sigma * dot(w->dQw, w->ATdy, n);
        m2 += 2.0 * dot(w->ATdy, w->dx, n);
        if (w->Q_is_diag) {
            for (size_t i = 0; i < n; ++i) w->tempv[i] = w->ATdy[i] * w->fact_M[i];
            m3 = dot(w->tempv, w->QATdy, n);
            m1 += w->sigma * dot(w->dQw, w->dQw, n);
        } else {
            m3 = (w->sigma * w->sigma) / (1.0 + w->sigma * w->lambda_max_Q) * dot(w->ATdy, w->QATdy, n);
            m1 += w->sigma * w->lambda_max_Q * dot(w->dw, w->dQw, n);
        }
    } else if (!w->Q_is_diag) {
        m1 = w->sigma * w->lambda_max_Q * dot(w->dw, w->dQw, n);
    }
    m2 += m1 > 0.0 ? m1 : 0.0;
    return sqrt((m2 > 0.0 ? m2 : 0.0) + (m3 > 0.0 ? m3 : 0.0));
}

/* control.c:231-282 */
static void restart_check(Restart *r, size_t iter, size_t check_iter, double sigma)
{
    double fact;
    if (check_iter == 0) return;
    r->flag = 0;
    if (r->first) {
        if (iter == check_iter) {
            r->first = 0;
            r->flag = 1;
            r->weighted_norm = r->current_gap;
            r->best_gap = r->current_gap;
            r->best_sigma = sigma;
        }
        return;
    }
    if (iter % check_iter != 0) return;
    if (r->current_gap < 0.0) r->current_gap = 1e-6;
    if (r->current_gap <= 0.2 * r->last_gap) {
        r->sufficient += 1;
        r->flag = 1;
    }
    if (r->current_gap <= 0.8 * r->last_gap && r->current_gap > r->save_gap) {
        r->necessary += 1;
        r->flag = 2;
    }
    fact = r->weighted_norm > 0.0 && r->current_gap / r->weighted_norm > 1e-1 ? 0.5 : 0.2;
    if ((double)r->inner >= fact * (double)iter) {
        r->long_restart += 1;
        r->flag = 3;
    }
    if (r->best_gap > r->current_gap) {
        r->best_gap = r->current_gap;
        r->best_sigma = sigma;
    }
    r->save_gap = r->current_gap;
}

/* control.c:284-305 */
static void restart_post(Restart *r, size_t iter, size_t check_iter, double gap)
{
    int end_of_block = check_iter > 0 && (iter + 1) % check_iter == 0;
    if (r->flag > 0 || end_of_block) {
        if (r->flag > 0) r->last_gap = gap;
        if (end_of_block) r->current_gap = gap;
    }
}

/* control.c:20-77 */
static double sigma_f(double x, double a, double b, double c, double d)
{
    return a * x + b / x + c * x * x / (1.0 + d * x);
}

static double golden(double a, double b, double c, double d)
{
    double lo = 1e-12, hi = 1e12;
    double g = (sqrt(5.0) - 1.0) / 2.0;
    double g2 = (3.0 - sqrt(5.0)) / 2.0;
    double x1 = lo + g2 * (hi - lo), x2 = lo + g * (hi - lo);
    double f1 = sigma_f(x1, a, b, c, d), f2 = sigma_f(x2, a, b, c, d);
    for (int it = 0; it < 256; ++it) {
        if (fabs(hi - lo) <= 1e-13 * fmax(1.0, fmax(fabs(lo), fabs(hi)))) break;
        if (f1 > f2) {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + g * (hi - lo);
            f2 = sigma_f(x2, a, b, c, d);
        } else {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = lo + g2 * (hi - lo);
            f1 = sigma_f(x1, a, b, c, d);
        }
    }
    return 0.5 * (lo + hi);
}

/* control.c:25-43 and 79-114 */
static double sigma_f_diag(double x, double a, double b, const double *dq, const double *c, const double *d,
                           double *tmp, size_t n)
{
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) {
        tmp[i] = d[i] / (1.0 + x * dq[i]);
        acc += c[i] * tmp[i];
    }
    return a * x + b / x + x * x * acc;
}

static double golden_diag(double a, double b, const double *dq, const double *c, const double *d, double *tmp,
                          size_t n)
{
    double lo = 1e-12, hi = 1e12;
    double phi = (sqrt(5.0) - 1.0) / 2.0;
    double x1 = hi - phi * (hi - lo), x2 = lo + phi * (hi - lo);
    double f1 = sigma_f_diag(x1, a, b, dq, c, d, tmp, n);
    double f2 = sigma_f_diag(x2, a, b, dq, c, d, tmp, n);
    size_t it = 0;
    while (fabs(hi - lo) > 1e-12 * fmax(1.0, fabs(lo)) && it < 200) {
        if (f1 > f2) {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + phi * (hi - lo);
            f2 = sigma_f_diag(x2, a, b, dq, c, d, tmp, n);
        } else {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = hi - phi * (hi - lo);
            f1 = sigma_f_diag(x1, a, b, dq, c, d, tmp, n);
        }
        it += 1;
    }
    return f1 < f2 ? x1 : x2;
}

/* control.c:307-411 */
static void sigma_update(Restart *r, Work *w, const Model *md, const Resid *res)
{
    size_t n = w->n, m = w->m;
    double old = w->sigma, a = 0.0, b, est, cand, fact, t1, kappa = 1.0, cterm = 0.0, dterm = 0.0;
    if (r->flag < 1) return;
    for (size_t i = 0; i < n; ++i) {
        w->dx[i] = w->x_bar[i] - w->last_x[i];
        w->dw[i] = w->w_bar[i] - w->last_w[i];
    }
    b = dot(w->dx, w->dx, n);
    spmv(&md->Q, w->dw, w->dQw);
    if (m > 0) {
        for (size_t i = 0; i < m; ++i) w->dy[i] = w->y_bar[i] - w->last_y[i];
        for (size_t i = 0; i < n; ++i) w->ATdy[i] = w->ATy_bar[i] - w->last_ATy[i];
        a = w->g.kind ? gram_quadratic(w) : w->lambda_max_A * dot(w->dy, w->dy, m);
        a -= 2.0 * dot(w->dQw, w->ATdy, n);
        spmv(&md->Q, w->ATdy, w->QATdy);
    }
    if (w->Q_is_diag) {
        a += dot(w->dQw, w->dQw, n);
    } else {
        a += w->lambda_max_Q * dot(w->dw, w->dQw, n);
        if (m > 0) {
            cterm = dot(w->ATdy, w->QATdy, n);
            dterm = w->lambda_max_Q;
        }
    }
    a = a > 1e-12 ? a : 1e-12;
    b = b > 1e-12 ? b : 1e-12;
    if (w->Q_is_diag && m > 0) {
        est = golden_diag(a, b, w->diag_Q, w->ATdy, w->QATdy, w->tempv, n);
    } else if (!w->Q_is_diag && m > 0) {
        cterm = cterm > 0.0 ? cterm : 0.0;
        dterm = dterm > 0.0 ? dterm : 0.0;
        est = golden(a, b, cterm, dterm);
    } else {
        est = sqrt(b / a);
    }
    est = clamp(est, 1e-12, 1e12);
    fact = (r->best_gap > 0.0 && isfinite(r->best_gap)) ? exp(-0.05 * (r->current_gap / r->best_gap)) : 1.0;
    cand = exp(fact * log(est) + (1.0 - fact) * log(r->best_sigma > 0.0 ? r->best_sigma : old));
    t1 = fmax(fmin(res->err_rd, res->err_rp), fmin(res->gap, r->current_gap));
    if (t1 <= 9e-10 && res->err_rp > 0.0) {
        double ratio = res->err_rd / res->err_rp;
        kappa = t1 > 5e-10 ? clamp(sqrt(ratio), 1e-2, 100.0) : clamp(ratio, 1e-2, 100.0);
    }
    w->sigma = clamp(kappa * cand, 1e-12, 1e12);
    if (w->Q_is_diag && fabs(w->sigma - old) > 1e-15) diag_factors(w);
}

/* control.c:413-457 */
static void restart_apply(Restart *r, Work *w, const Model *md)
{
    size_t n = w->n, m = w->m;
    if (r->flag <= 0) return;
    for (size_t i = 0; i < n; ++i) {
        w->x[i] = w->x_bar[i];
        w->w[i] = w->w_bar[i];
        w->last_x[i] = w->x_bar[i];
        w->last_w[i] = w->w_bar[i];
    }
    if (m > 0) {
        for (size_t i = 0; i < m; ++i) {
            w->y[i] = w->y_bar[i];
            w->last_y[i] = w->y_bar[i];
        }
        spmv(&md->AT, w->y_bar, w->ATy_bar);
        for (size_t i = 0; i < n; ++i) {
            w->last_ATy[i] = w->ATy_bar[i];
            w->ATy[i] = w->ATy_bar[i];
        }
    }
    if (w->noC) {
        for (size_t i = 0; i < n; ++i) {
            w->Qw[i] = w->Qw_bar[i];
            w->last_Qw[i] = w->Qw_bar[i];
        }
    }
    r->last_gap = r->current_gap;
    r->save_gap = INFINITY;
    r->times += 1;
    r->inner = 0;
}

/* ------------------------------------------------------------------------ */
/* Solve loop: src/shared/solver_backend.c:119-569                           */

struct OrSession {
    Model md;
    Scaling sc;
    Work w;
    OrParams p;
    Restart rs;
    Resid res;
    size_t iter;
    int started, finished, status, first4, first6;
    double start;
    OrResults out;
};

static size_t print_step(size_t iter)
{
    size_t step = 10;
    if (iter == 0) return step;
    while (iter / step >= 100 && step <= SIZE_MAX / 10) step *= 10;
    return step;
}

/* solver_backend.c:134-149 */
static int print_now(const OrParams *p, size_t iter, double el)
{
    if (p->print_frequency == -1 && (iter % print_step(iter) == 0 || iter == p->max_iter || el > p->time_limit))
        return 1;
    if (p->print_frequency > 0 &&
        (iter % (size_t)p->print_frequency == 0 || iter == p->max_iter || el > p->time_limit))
        return 1;
    return 0;
}

/* solver_backend.c:151-171 */
static int check_next(const OrParams *p, const Restart *r, size_t next)
{
    int chk = 0;
    if (p->check_iter > 0 && next % p->check_iter == 0) chk = 1;
    if (r->flag > 0) chk = 1;
    if (p->print_frequency == -1 && next % print_step(next) == 0) chk = 1;
    if (p->print_frequency > 0 && next % (size_t)p->print_frequency == 0) chk = 1;
    return chk;
}

static void milestones(OrSession *s, size_t iter, double el)
{
    if (s->first4 && s->res.kkt < 1e-4) {
        s->out.iter_4 = iter;
        s->out.time_4 = el;
        s->first4 = 0;
    }
    if (s->first6 && s->res.kkt < 1e-6) {
        s->out.iter_6 = iter;
        s->out.time_6 = el;
        s->first6 = 0;
    }
}

static double elapsed(const OrSession *s)
{
    double now = wall_seconds();
    if (now <= 0.0 || s->start <= 0.0) return 0.0;
    return now - s->start;
}

/* solver_backend.c:192-208 */
static int termination(const OrSession *s, size_t iter, double el)
{
    if (s->res.kkt < s->p.stoptol) return OR_OPTIMAL;
    if (iter == s->p.max_iter) return OR_MAX_ITER;
    if (el > s->p.time_limit) return OR_TIME_LIMIT;
    return OR_CONTINUE;
}

static void loop_begin(OrSession *s)
{
    restart_reset(&s->rs);
    s->start = wall_seconds();
    memset(&s->out, 0, sizeof(s->out));
    s->first4 = s->first6 = 1;
    residuals(&s->w, &s->md, &s->sc, &s->res, 0);
    milestones(s, 0, elapsed(s));
    s->rs.best_sigma = s->w.sigma;
    s->iter = 0;
    s->started = 1;
}

/* One body of the loop at solver_backend.c:432-513. */
static int loop_body(OrSession *s)
{
    size_t iter = s->iter;
    double el = elapsed(s);
    int st;
    if (iter > 0) {
        int pr = print_now(&s->p, iter, el);
        if ((s->p.check_iter > 0 && iter % s->p.check_iter == 0) || pr) {
            residuals(&s->w, &s->md, &s->sc, &s->res, iter);
            milestones(s, iter, elapsed(s));
        } else {
            s->res.updated = 0;
        }
    }
    s->status = termination(s, iter, el);
    restart_check(&s->rs, iter, s->p.check_iter, s->w.sigma);
    sigma_update(&s->rs, &s->w, &s->md, &s->res);
    restart_apply(&s->rs, &s->w, &s->md);
    if (s->status != OR_CONTINUE) {
        s->finished = 1;
        return OR_OK;
    }
    s->w.to_check = check_next(&s->p, &s->rs, iter + 1);
    if ((st = iteration_step(&s->w, &s->md, s->rs.inner)) != OR_OK) {
        s->finished = 1;
        s->status = st;
        return st;
    }
    if (s->rs.flag > 0 || (s->p.check_iter > 0 && (iter + 1) % s->p.check_iter == 0)) {
        double g = m_norm(&s->w, &s->md);
        if (!isfinite(g)) {
            s->finished = 1;
            s->status = OR_INVALID;
            return OR_INVALID;
        }
        restart_post(&s->rs, iter, s->p.check_iter, g);
    }
    s->rs.inner += 1;
    s->iter = iter + 1;
    return OR_OK;
}

void or_params_defaults(OrParams *p)
{
    p->stoptol = 1e-8;
    p->max_iter = (size_t)2147483647;
    p->time_limit = 3600.0;
    p->check_iter = 100;
    p->print_frequency = -1;
    p->use_scaling = 1;
    p->lambda_A_given = 0.0;
    p->lambda_Q_given = 0.0;
    p->gram_mode = 1;
}

int or_session_create(size_t n, size_t m,
                      size_t q_nnz, const size_t *q_ptr, const size_t *q_ind, const double *q_val,
                      size_t a_nnz, const size_t *a_ptr, const size_t *a_ind, const double *a_val,
                      const double *c, const double *AL, const double *AU,
                      const double *l, const double *u, double obj_constant,
                      size_t num_var_cones, const size_t *var_offsets,
                      size_t num_row_cones, const size_t *row_offsets, const double *soc_rhs,
                      const OrParams *params, OrSession **out)
{
    OrSession *s;
    int st;
    if (out == NULL || params == NULL) return OR_INVALID;
    *out = NULL;
    s = (OrSession *)calloc(1, sizeof(OrSession));
    if (!s) return OR_ALLOC;
    s->p = *params;
    st = model_build(&s->md, n, m, q_nnz, q_ptr, q_ind, q_val, a_nnz, a_ptr, a_ind, a_val, c, AL, AU, l, u,
                     obj_constant, num_var_cones, var_offsets, num_row_cones, row_offsets, soc_rhs);
    /* solver.cu:145-151: AUTO becomes REQUEST when the built model passes
     * the structural rule, else OFF */
    if (st == OR_OK && (s->p.gram_mode < 0 || s->p.gram_mode > 2)) st = OR_INVALID;
    if (st == OR_OK && s->p.gram_mode == 1) s->p.gram_mode = gram_auto_eligible(&s->md) ? 2 : 0;
    if (st == OR_OK) st = scale_model(&s->md, params->use_scaling, &s->sc);
    if (st == OR_OK) st = work_setup(&s->w, &s->md, &s->p);
    if (st != OR_OK) {
        or_session_destroy(s);
        return st;
    }
    *out = s;
    return OR_OK;
}

void or_session_destroy(OrSession *s)
{
    if (!s) return;
    model_free(&s->md);
    free(s->sc.row_norm);
    free(s->sc.col_norm);
    work_free(&s->w);
    free(s);
}

int or_session_advance(OrSession *s, size_t bodies, int *finished)
{
    int st = OR_OK;
    if (!s) return OR_INVALID;
    if (!s->started) loop_begin(s);
    for (size_t k = 0; k < bodies && !s->finished; ++k) {
        st = loop_body(s);
        if (st != OR_OK) break;
    }
    if (finished) *finished = s->finished;
    return st;
}

/* solver.c:67-106 and solver_backend.c:289-336 */
int or_session_results(OrSession *s, OrResults *res, double *x, double *y, double *z, double *w)
{
    double el = elapsed(s);
    if (!s || !res) return OR_INVALID;
    *res = s->out;
    res->status = s->finished ? s->status : OR_CONTINUE;
    res->iter = s->iter;
    res->time = el;
    res->iter_4 = s->out.iter_4 == 0 ? s->iter : s->out.iter_4;
    res->iter_6 = s->out.iter_6 == 0 ? s->iter : s->out.iter_6;
    res->time_4 = s->out.time_4 == 0.0 ? el : s->out.time_4;
    res->time_6 = s->out.time_6 == 0.0 ? el : s->out.time_6;
    res->primal_obj = s->res.pobj;
    res->residuals = s->res.kkt;
    res->gap = s->res.gap;
    res->err_rp = s->res.err_rp;
    res->err_rd = s->res.err_rd;
    res->dual_obj = s->res.dobj;
    for (size_t i = 0; i < s->w.n; ++i) {
        double cn = s->sc.col_norm[i];
        if (x) x[i] = s->sc.b_scale * (s->w.x_bar[i] / cn);
        if (w) w[i] = s->sc.b_scale * (s->w.w_bar[i] / cn);
        if (z) z[i] = s->sc.c_scale * (s->w.z_bar[i] * cn);
    }
    for (size_t i = 0; i < s->w.m; ++i)
        if (y) y[i] = s->sc.c_scale * (s->w.y_bar[i] / s->sc.row_norm[i]);
    return OR_OK;
}

int or_session_solve(OrSession *s, OrResults *res, double *x, double *y, double *z, double *w)
{
    int fin = 0, st = or_session_advance(s, (size_t)-1, &fin);
    if (st != OR_OK) return st;
    return or_session_results(s, res, x, y, z, w);
}

size_t or_session_n(const OrSession *s) { return s ? s->w.n : 0; }
size_t or_session_m(const OrSession *s) { return s ? s->w.m : 0; }

size_t or_session_read(const OrSession *s, const char *name, double *out, size_t cap)
{
    const double *src = NULL;
    size_t len = 0;
    if (!s || !name) return 0;
#define OR_TRY_N(v) if (!src && strcmp(name, #v) == 0) { src = s->w.v; len = s->w.n; }
#define OR_TRY_M(v) if (!src && strcmp(name, #v) == 0) { src = s->w.v; len = s->w.m; }
    OR_WORK_N_VECTORS(OR_TRY_N)
    OR_WORK_M_VECTORS(OR_TRY_M)
#undef OR_TRY_N
#undef OR_TRY_M
    if (!src && strcmp(name, "c") == 0) { src = s->md.c; len = s->md.n; }
    if (!src && strcmp(name, "l") == 0) { src = s->md.l; len = s->md.n; }
    if (!src && strcmp(name, "u") == 0) { src = s->md.u; len = s->md.n; }
    if (!src && strcmp(name, "AL") == 0) { src = s->md.AL; len = s->md.m; }
    if (!src && strcmp(name, "AU") == 0) { src = s->md.AU; len = s->md.m; }
    if (!src && strcmp(name, "soc_rhs") == 0 && s->md.soc_rhs) { src = s->md.soc_rhs; len = s->md.m; }
    if (!src && strcmp(name, "row_norm") == 0) { src = s->sc.row_norm; len = s->md.m; }
    if (!src && strcmp(name, "col_norm") == 0) { src = s->sc.col_norm; len = s->md.n; }
    if (!src && s->w.g.kind && strcmp(name, "gram_G") == 0) { src = s->w.g.G; len = s->w.g.e * s->w.g.e; }
    if (!src && s->w.g.kind && strcmp(name, "gram_r") == 0) { src = s->w.g.r; len = s->w.g.e; }
    if (!src && s->w.g.kind && strcmp(name, "gram_step") == 0) { src = s->w.g.step; len = s->w.g.e; }
    if (!src && s->w.g.kind && strcmp(name, "gram_b") == 0) { src = s->w.g.b; len = s->w.g.e; }
    if (!src) return 0;
    if (out) memcpy(out, src, (len < cap ? len : cap) * sizeof(double));
    return len;
}

double or_session_scalar(const OrSession *s, const char *name)
{
    if (!s || !name) return NAN;
    if (!strcmp(name, "sigma")) return s->w.sigma;
    if (!strcmp(name, "lambda_max_A")) return s->w.lambda_max_A;
    if (!strcmp(name, "lambda_max_Q")) return s->w.lambda_max_Q;
    if (!strcmp(name, "power_steps_A")) return (double)s->w.power_steps_A;
    if (!strcmp(name, "power_steps_Q")) return (double)s->w.power_steps_Q;
    if (!strcmp(name, "Q_is_diag")) return s->w.Q_is_diag;
    if (!strcmp(name, "noC")) return s->w.noC;
    if (!strcmp(name, "iter")) return (double)s->iter;
    if (!strcmp(name, "inner")) return (double)s->rs.inner;
    if (!strcmp(name, "restart_times")) return (double)s->rs.times;
    if (!strcmp(name, "norm_b_org")) return s->sc.norm_b_org;
    if (!strcmp(name, "norm_c_org")) return s->sc.norm_c_org;
    if (!strcmp(name, "norm_b")) return s->sc.norm_b;
    if (!strcmp(name, "norm_c")) return s->sc.norm_c;
    if (!strcmp(name, "kkt")) return s->res.kkt;
    if (!strcmp(name, "err_rp")) return s->res.err_rp;
    if (!strcmp(name, "err_rd")) return s->res.err_rd;
    if (!strcmp(name, "gap")) return s->res.gap;
    if (!strcmp(name, "pobj")) return s->res.pobj;
    if (!strcmp(name, "dobj")) return s->res.dobj;
    if (!strcmp(name, "current_gap")) return s->rs.current_gap;
    if (!strcmp(name, "last_gap")) return s->rs.last_gap;
    if (!strcmp(name, "gram_active")) return s->w.g.kind;
    if (!strcmp(name, "gram_rows")) return (double)s->w.g.e;
    if (!strcmp(name, "gram_beta")) return s->w.g.beta;
    if (!strcmp(name, "gram_nnz")) return (double)s->w.g.gnnz;
    if (!strcmp(name, "gram_backward")) return s->w.g.backward;
    if (!strcmp(name, "gram_mode")) return s->p.gram_mode;
    return NAN;
}

size_t or_session_matrix(const OrSession *s, const char *name, size_t *ptr, size_t *ind, double *val)
{
    const Csr *a = NULL;
    if (!s || !name) return 0;
    if (!strcmp(name, "Q")) a = &s->md.Q;
    else if (!strcmp(name, "A")) a = &s->md.A;
    else if (!strcmp(name, "AT")) a = &s->md.AT;
    if (!a) return 0;
    if (ptr) memcpy(ptr, a->ptr, (a->nrow + 1) * sizeof(size_t));
    if (ind && a->nnz) memcpy(ind, a->ind, a->nnz * sizeof(size_t));
    if (val && a->nnz) memcpy(val, a->val, a->nnz * sizeof(double));
    return a->nnz;
}

int or_preprocess(size_t n, size_t m,
                  size_t q_nnz, const size_t *q_ptr, const size_t *q_ind, const double *q_val,
                  size_t a_nnz, const size_t *a_ptr, const size_t *a_ind, const double *a_val,
                  const double *AL, const double *AU,
                  size_t *q_nnz_out, size_t *q_ptr_out, size_t *q_ind_out, double *q_val_out,
                  size_t *m_out, size_t *a_nnz_out, size_t *a_ptr_out, size_t *a_ind_out, double *a_val_out)
{
    Model md;
    int st;
    double *zeros = (double *)calloc(n ? n : 1, sizeof(double));
    double *lo = (double *)malloc((n ? n : 1) * sizeof(double));
    double *hi = (double *)malloc((n ? n : 1) * sizeof(double));
    for (size_t i = 0; i < n; ++i) {
        lo[i] = -INFINITY;
        hi[i] = INFINITY;
    }
    st = model_build(&md, n, m, q_nnz, q_ptr, q_ind, q_val, a_nnz, a_ptr, a_ind, a_val, zeros, AL, AU, lo, hi,
                     0.0, 0, NULL, 0, NULL, NULL);
    if (st == OR_OK) {
        if (q_nnz_out) *q_nnz_out = md.Q.nnz;
        if (m_out) *m_out = md.m;
        if (a_nnz_out) *a_nnz_out = md.A.nnz;
        if (q_ptr_out) memcpy(q_ptr_out, md.Q.ptr, (md.Q.nrow + 1) * sizeof(size_t));
        if (q_ind_out && md.Q.nnz) memcpy(q_ind_out, md.Q.ind, md.Q.nnz * sizeof(size_t));
        if (q_val_out && md.Q.nnz) memcpy(q_val_out, md.Q.val, md.Q.nnz * sizeof(double));
        if (a_ptr_out) memcpy(a_ptr_out, md.A.ptr, (md.A.nrow + 1) * sizeof(size_t));
        if (a_ind_out && md.A.nnz) memcpy(a_ind_out, md.A.ind, md.A.nnz * sizeof(size_t));
        if (a_val_out && md.A.nnz) memcpy(a_val_out, md.A.val, md.A.nnz * sizeof(double));
    }
    model_free(&md);
    free(zeros);
    free(lo);
    free(hi);
    return st;
}

/* Direct access for closed-form unit checks (the reference's own unit tests
 * drive workspace_allocate / iteration_step / residuals_compute directly). */
int or_session_write(OrSession *s, const char *name, const double *in, size_t len)
{
    double *dst = NULL;
    size_t n;
    if (!s || !name || !in) return OR_INVALID;
#define OR_W_N(v) if (!dst && strcmp(name, #v) == 0) { dst = s->w.v; n = s->w.n; }
#define OR_W_M(v) if (!dst && strcmp(name, #v) == 0) { dst = s->w.v; n = s->w.m; }
    OR_WORK_N_VECTORS(OR_W_N)
    OR_WORK_M_VECTORS(OR_W_M)
#undef OR_W_N
#undef OR_W_M
    if (!dst || len > n) return OR_INVALID;
    memcpy(dst, in, len * sizeof(double));
    return OR_OK;
}

int or_session_step(OrSession *s, size_t inner, int to_check)
{
    if (!s) return OR_INVALID;
    s->w.to_check = to_check;
    return iteration_step(&s->w, &s->md, inner);
}

int or_session_residuals(OrSession *s, size_t iter, double *out6)
{
    if (!s || !out6) return OR_INVALID;
    residuals(&s->w, &s->md, &s->sc, &s->res, iter);
    out6[0] = s->res.pobj;
    out6[1] = s->res.dobj;
    out6[2] = s->res.err_rd;
    out6[3] = s->res.err_rp;
    out6[4] = s->res.gap;
    out6[5] = s->res.kkt;
    return OR_OK;
}

double or_session_m_norm(OrSession *s)
{
    return s ? m_norm(&s->w, &s->md) : NAN;
}

double or_session_gram_quadratic(OrSession *s)
{
    return s && s->w.g.kind ? gram_quadratic(&s->w) : NAN;
}

size_t or_session_gram_map(const OrSession *s, long *out, size_t cap)
{
    if (!s || !s->w.g.kind) return 0;
    if (out) memcpy(out, s->w.g.map, (s->w.m < cap ? s->w.m : cap) * sizeof(long));
    return s->w.m;
}
