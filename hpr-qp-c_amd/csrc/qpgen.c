/*
 * Seeded synthetic QP generators for the benchmark workloads.
 *
 * Random sparse convex QP ("plain" workloads):
 *   min 1/2 x'Qx + c'x  s.t.  A_eq x = b, l_A <= A_in x <= u_A, -box <= x <= box
 *   A: a_row_nnz distinct uniform columns per row, values N(0,1); `heavy_rows`
 *      rows spread evenly over A carry heavy_nnz entries instead;
 *   Q = B' D B + 1e-2 I, B (b_rows x n) at b_row_nnz per row N(0,1),
 *      D = diag U(0.1, 1);
 *   c ~ N(0,1); x0 ~ U(-1,1); b = A_eq x0; l/u = A_in x0 -/+ U(0,1).
 * Second-order-cone QP: a box prefix then `cones` Lorentz cones of `cone_dim`
 *   variables (x_0 >= ||x_1:||), equality rows at a_row_nnz, Q diagonal U(0,1),
 *   b = A x0 with x0 strictly inside the cones.
 *
 * PRNG: SplitMix64; U[0,1) from the top 53 bits; N(0,1) by Box-Muller (both
 * outputs used). Everything is FP64 and deterministic for a given seed.
 */
#include "qpgen.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct Rng {
    uint64_t s;
    int has_spare;
    double spare;
} Rng;

static uint64_t rng_next(Rng *r)
{
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double rng_uniform(Rng *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

static double rng_normal(Rng *r)
{
    double u1, u2, rad;
    if (r->has_spare) {
        r->has_spare = 0;
        return r->spare;
    }
    do {
        u1 = rng_uniform(r);
    } while (u1 <= 0.0);
    u2 = rng_uniform(r);
    rad = sqrt(-2.0 * log(u1));
    r->spare = rad * sin(6.283185307179586 * u2);
    r->has_spare = 1;
    return rad * cos(6.283185307179586 * u2);
}

static int cmp_size(const void *a, const void *b)
{
    size_t x = *(const size_t *)a, y = *(const size_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* k distinct columns in [0, n), ascending: draw, sort, dedupe, redraw. */
static size_t sample_columns(Rng *r, size_t n, size_t k, size_t *out)
{
    size_t have = 0;
    if (k > n) k = n;
    if (k == 0) return 0;
    while (have < k) {
        size_t u = 1;
        for (size_t t = have; t < k; ++t) out[t] = (size_t)(rng_next(r) % n);
        qsort(out, k, sizeof(size_t), cmp_size);
        for (size_t t = 1; t < k; ++t)
            if (out[t] != out[u - 1]) out[u++] = out[t];
        have = u;
    }
    return k;
}

void qpgen_free(QPGenProblem *p)
{
    if (!p) return;
    free(p->q_ptr);
    free(p->q_ind);
    free(p->q_val);
    free(p->a_ptr);
    free(p->a_ind);
    free(p->a_val);
    free(p->c);
    free(p->AL);
    free(p->AU);
    free(p->l);
    free(p->u);
    free(p->x0);
    free(p->var_offsets);
    free(p->row_offsets);
    free(p->soc_rhs);
    memset(p, 0, sizeof(*p));
}

/* Sparse A with `rows` rows; heavy rows at evenly spaced positions. */
static int gen_matrix_a(Rng *r, size_t rows, size_t n, size_t k, size_t heavy_rows, size_t heavy_nnz,
                        size_t **ptr_out, size_t **ind_out, double **val_out)
{
    size_t *ptr = (size_t *)calloc(rows + 1, sizeof(size_t));
    size_t cap, nnz = 0, next_heavy = 0;
    size_t *ind;
    double *val;
    size_t *scratch;
    size_t maxk = heavy_nnz > k ? heavy_nnz : k;
    if (!ptr) return -1;
    cap = rows * (k < n ? k : n) + heavy_rows * (heavy_nnz < n ? heavy_nnz : n) + 1;
    ind = (size_t *)malloc(cap * sizeof(size_t));
    val = (double *)malloc(cap * sizeof(double));
    scratch = (size_t *)malloc((maxk + 1) * sizeof(size_t));
    if (!ind || !val || !scratch) return -1;
    for (size_t i = 0; i < rows; ++i) {
        size_t want = k, got;
        if (next_heavy < heavy_rows) {
            size_t pos = (size_t)(((double)next_heavy + 0.5) * (double)rows / (double)heavy_rows);
            if (pos == i) {
                want = heavy_nnz;
                ++next_heavy;
            }
        }
        got = sample_columns(r, n, want, scratch);
        for (size_t t = 0; t < got; ++t) {
            ind[nnz] = scratch[t];
            val[nnz] = rng_normal(r);
            ++nnz;
        }
        ptr[i + 1] = nnz;
    }
    free(scratch);
    *ptr_out = ptr;
    *ind_out = ind;
    *val_out = val;
    return 0;
}

typedef struct Trip {
    size_t row, col;
    double val;
} Trip;

/* Q = B' D B + shift I, full symmetric CSR, columns ascending. Contributions
 * to one (i, j) are summed in increasing B-row order; the shift goes last. */
static int gen_gram(Rng *r, size_t n, size_t b_rows, size_t b_k, double shift, QPGenProblem *p)
{
    size_t k = b_k < n ? b_k : n;
    size_t ntrip = b_rows * k * k + n, cnt = 0, out = 0;
    Trip *t = (Trip *)malloc(ntrip * sizeof(Trip));
    Trip *sorted = (Trip *)malloc(ntrip * sizeof(Trip));
    size_t *start = (size_t *)calloc(n + 1, sizeof(size_t));
    size_t *cols = (size_t *)malloc((k + 1) * sizeof(size_t));
    double *vals = (double *)malloc((k + 1) * sizeof(double));
    if (!t || !sorted || !start || !cols || !vals) return -1;
    for (size_t b = 0; b < b_rows; ++b) {
        size_t got = sample_columns(r, n, k, cols);
        double d;
        for (size_t a = 0; a < got; ++a) vals[a] = rng_normal(r);
        d = 0.1 + 0.9 * rng_uniform(r);
        for (size_t a = 0; a < got; ++a)
            for (size_t c = 0; c < got; ++c) {
                t[cnt].row = cols[a];
                t[cnt].col = cols[c];
                t[cnt].val = d * vals[a] * vals[c];
                ++cnt;
            }
    }
    for (size_t i = 0; i < n; ++i) {
        t[cnt].row = i;
        t[cnt].col = i;
        t[cnt].val = shift;
        ++cnt;
    }
    for (size_t e = 0; e < cnt; ++e) start[t[e].row + 1] += 1;
    for (size_t i = 0; i < n; ++i) start[i + 1] += start[i];
    {
        size_t *fill = (size_t *)malloc((n + 1) * sizeof(size_t));
        memcpy(fill, start, (n + 1) * sizeof(size_t));
        for (size_t e = 0; e < cnt; ++e) sorted[fill[t[e].row]++] = t[e];
        free(fill);
    }
    free(t);
    p->q_ptr = (size_t *)calloc(n + 1, sizeof(size_t));
    p->q_ind = (size_t *)malloc(cnt * sizeof(size_t));
    p->q_val = (double *)malloc(cnt * sizeof(double));
    if (!p->q_ptr || !p->q_ind || !p->q_val) return -1;
    for (size_t i = 0; i < n; ++i) {
        size_t lo = start[i], hi = start[i + 1];
        /* stable insertion sort by column: rows are short */
        for (size_t a = lo + 1; a < hi; ++a) {
            Trip key = sorted[a];
            size_t b = a;
            while (b > lo && sorted[b - 1].col > key.col) {
                sorted[b] = sorted[b - 1];
                --b;
            }
            sorted[b] = key;
        }
        for (size_t a = lo; a < hi; ++a) {
            if (out > p->q_ptr[i] && p->q_ind[out - 1] == sorted[a].col) {
                p->q_val[out - 1] += sorted[a].val;
            } else {
                p->q_ind[out] = sorted[a].col;
                p->q_val[out] = sorted[a].val;
                ++out;
            }
        }
        p->q_ptr[i + 1] = out;
    }
    p->q_nnz = out;
    free(sorted);
    free(start);
    free(cols);
    free(vals);
    return 0;
}

int qpgen_random_qp(uint64_t seed, size_t n, size_t m_eq, size_t m_in, size_t a_row_nnz, size_t heavy_rows,
                    size_t heavy_nnz, size_t b_rows, size_t b_row_nnz, double box, QPGenProblem *p)
{
    Rng r = {seed, 0, 0.0};
    size_t m = m_eq + m_in;
    double *ax;
    if (!p || n == 0) return -1;
    memset(p, 0, sizeof(*p));
    p->n = n;
    p->m = m;
    p->m_eq = m_eq;
    if (gen_matrix_a(&r, m, n, a_row_nnz, heavy_rows, heavy_nnz, &p->a_ptr, &p->a_ind, &p->a_val) != 0) return -1;
    p->a_nnz = p->a_ptr[m];
    if (gen_gram(&r, n, b_rows, b_row_nnz, 1e-2, p) != 0) return -1;
    p->c = (double *)malloc(n * sizeof(double));
    p->x0 = (double *)malloc(n * sizeof(double));
    p->l = (double *)malloc(n * sizeof(double));
    p->u = (double *)malloc(n * sizeof(double));
    p->AL = (double *)malloc((m ? m : 1) * sizeof(double));
    p->AU = (double *)malloc((m ? m : 1) * sizeof(double));
    ax = (double *)malloc((m ? m : 1) * sizeof(double));
    if (!p->c || !p->x0 || !p->l || !p->u || !p->AL || !p->AU || !ax) return -1;
    for (size_t j = 0; j < n; ++j) p->c[j] = rng_normal(&r);
    for (size_t j = 0; j < n; ++j) p->x0[j] = -1.0 + 2.0 * rng_uniform(&r);
    for (size_t j = 0; j < n; ++j) {
        p->l[j] = -box;
        p->u[j] = box;
    }
    for (size_t i = 0; i < m; ++i) {
        double acc = 0.0;
        for (size_t k = p->a_ptr[i]; k < p->a_ptr[i + 1]; ++k) acc += p->a_val[k] * p->x0[p->a_ind[k]];
        ax[i] = acc;
    }
    for (size_t i = 0; i < m; ++i) {
        if (i < m_eq) {
            p->AL[i] = ax[i];
            p->AU[i] = ax[i];
        } else {
            p->AL[i] = ax[i] - rng_uniform(&r);
            p->AU[i] = ax[i] + rng_uniform(&r);
        }
    }
    free(ax);
    return 0;
}

int qpgen_soc_qp(uint64_t seed, size_t n_box, size_t cones, size_t cone_dim, size_t m, size_t a_row_nnz,
                 double box, QPGenProblem *p)
{
    Rng r = {seed, 0, 0.0};
    size_t n = n_box + cones * cone_dim;
    if (!p || n == 0 || cone_dim < 2) return -1;
    memset(p, 0, sizeof(*p));
    p->n = n;
    p->m = m;
    p->m_eq = m;
    if (gen_matrix_a(&r, m, n, a_row_nnz, 0, 0, &p->a_ptr, &p->a_ind, &p->a_val) != 0) return -1;
    p->a_nnz = p->a_ptr[m];
    p->q_ptr = (size_t *)calloc(n + 1, sizeof(size_t));
    p->q_ind = (size_t *)malloc(n * sizeof(size_t));
    p->q_val = (double *)malloc(n * sizeof(double));
    p->c = (double *)malloc(n * sizeof(double));
    p->x0 = (double *)malloc(n * sizeof(double));
    p->l = (double *)malloc(n * sizeof(double));
    p->u = (double *)malloc(n * sizeof(double));
    p->AL = (double *)malloc((m ? m : 1) * sizeof(double));
    p->AU = (double *)malloc((m ? m : 1) * sizeof(double));
    if (!p->q_ptr || !p->q_ind || !p->q_val || !p->c || !p->x0 || !p->l || !p->u || !p->AL || !p->AU) return -1;
    for (size_t j = 0; j < n; ++j) {
        p->q_ind[j] = j;
        p->q_val[j] = rng_uniform(&r);
        p->q_ptr[j + 1] = j + 1;
    }
    p->q_nnz = n;
    for (size_t j = 0; j < n; ++j) p->c[j] = rng_normal(&r);
    for (size_t j = 0; j < n_box; ++j) {
        p->x0[j] = -1.0 + 2.0 * rng_uniform(&r);
        p->l[j] = -box;
        p->u[j] = box;
    }
    for (size_t k = 0; k < cones; ++k) {
        size_t b = n_box + k * cone_dim;
        double ss = 0.0;
        for (size_t j = 1; j < cone_dim; ++j) {
            p->x0[b + j] = -1.0 + 2.0 * rng_uniform(&r);
            ss += p->x0[b + j] * p->x0[b + j];
        }
        p->x0[b] = sqrt(ss) + 0.5 + 0.5 * rng_uniform(&r);
        for (size_t j = 0; j < cone_dim; ++j) {
            p->l[b + j] = -INFINITY;
            p->u[b + j] = INFINITY;
        }
    }
    for (size_t i = 0; i < m; ++i) {
        double acc = 0.0;
        for (size_t k = p->a_ptr[i]; k < p->a_ptr[i + 1]; ++k) acc += p->a_val[k] * p->x0[p->a_ind[k]];
        p->AL[i] = acc;
        p->AU[i] = acc;
    }
    if (cones) {
        p->num_var_cones = cones;
        p->var_offsets = (size_t *)malloc((cones + 1) * sizeof(size_t));
        for (size_t k = 0; k <= cones; ++k) p->var_offsets[k] = n_box + k * cone_dim;
    }
    return 0;
}
