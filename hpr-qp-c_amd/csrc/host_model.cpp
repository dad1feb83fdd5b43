// Host-side model construction for the HPR-QP C API: CSR helpers, model
// ownership, build_from_QAbc[_soc] with the reference's preprocessing
// (validation, Q <- (Q + Q')/2 with drop tolerance 1e-15, deletion of empty and
// free rows), parameter defaults and result ownership.
//
// Semantics follow src/shared/build.c:40-621, src/shared/sparse.c:1-181 and
// src/shared/solver.c:133-177; the code is this library's own.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <thread>
#include <vector>

#include "device_build.h"
#include "hprqp.h"

namespace {

// ---------------------------------------------------------- host threading
// Large models are built on worker threads; every parallel step below
// produces exactly the sequential result (fixed partitions, ordered merges).

size_t host_workers(size_t work, size_t grain)
{
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    return std::max<size_t>(1, std::min<size_t>({16, hw, work / grain}));
}

// f(t) for t in [0, T), t = 0 on the calling thread.
template <class F>
void run_workers(size_t T, const F &f)
{
    std::vector<std::thread> pool;
    for (size_t t = 1; t < T; ++t) pool.emplace_back([&f, t] { f(t); });
    f(0);
    for (std::thread &th : pool) th.join();
}

// Row ranges of about equal entry counts: rows [cut[t], cut[t+1]).
std::vector<size_t> row_cuts(size_t nrow, const size_t *rp, size_t T)
{
    std::vector<size_t> cut(T + 1, nrow);
    cut[0] = 0;
    for (size_t t = 1; t < T; ++t) {
        const size_t target = rp[nrow] / T * t;
        cut[t] = (size_t)(std::lower_bound(rp, rp + nrow + 1, target) - rp);
        cut[t] = std::max(cut[t], cut[t - 1]);
    }
    return cut;
}

void parallel_copy(void *dst, const void *src, size_t bytes)
{
    const size_t T = host_workers(bytes, (size_t)16 << 20);
    const size_t chunk = (bytes + T - 1) / T;
    run_workers(T, [&](size_t t) {
        const size_t b = std::min(bytes, t * chunk), e = std::min(bytes, b + chunk);
        if (e > b) std::memcpy(static_cast<char *>(dst) + b, static_cast<const char *>(src) + b, e - b);
    });
}

// Stable counting-sort transpose of a CSR pattern: tptr (ncol + 1) receives
// the transposed row pointers and emit(dst, r, k) is called for every entry
// k of row r with its position dst in the transpose, entries of one column
// in ascending row order. Worker threads take contiguous row ranges with a
// column histogram each; within a column, range t's entries are placed
// before range t + 1's, which is the sequential order.
template <class Emit>
void transpose_scatter(size_t nrow, size_t ncol, const size_t *rp, const size_t *ci, size_t *tptr, const Emit &emit)
{
    const size_t nnz = rp[nrow];
    size_t T = host_workers(nnz, (size_t)1 << 20);
    // per-thread 32-bit histograms: bounded memory and no overflow
    if (nnz >= ((size_t)1 << 32)) T = 1;
    T = std::min(T, std::max<size_t>(1, ((size_t)1 << 28) / std::max<size_t>(ncol, 1)));
    std::fill(tptr, tptr + ncol + 1, 0);
    if (T == 1) {
        for (size_t k = 0; k < nnz; ++k) tptr[ci[k] + 1] += 1;
        for (size_t j = 0; j < ncol; ++j) tptr[j + 1] += tptr[j];
        std::vector<size_t> next(tptr, tptr + ncol);
        for (size_t r = 0; r < nrow; ++r)
            for (size_t k = rp[r]; k < rp[r + 1]; ++k) emit(next[ci[k]]++, r, k);
        return;
    }
    const std::vector<size_t> cut = row_cuts(nrow, rp, T);
    std::vector<std::vector<uint32_t>> hist(T);
    run_workers(T, [&](size_t t) {
        hist[t].assign(ncol, 0);
        for (size_t k = rp[cut[t]]; k < rp[cut[t + 1]]; ++k) hist[t][ci[k]] += 1;
    });
    // column totals, then per-range start offsets within each column
    const size_t cchunk = (ncol + T - 1) / T;
    run_workers(T, [&](size_t t) {
        const size_t b = std::min(ncol, t * cchunk), e = std::min(ncol, b + cchunk);
        for (size_t j = b; j < e; ++j) {
            size_t sum = 0;
            for (size_t u = 0; u < T; ++u) sum += hist[u][j];
            tptr[j + 1] = sum;
        }
    });
    for (size_t j = 0; j < ncol; ++j) tptr[j + 1] += tptr[j];
    run_workers(T, [&](size_t t) {
        const size_t b = std::min(ncol, t * cchunk), e = std::min(ncol, b + cchunk);
        for (size_t j = b; j < e; ++j) {
            size_t run = tptr[j];
            for (size_t u = 0; u < T; ++u) {
                const size_t c = hist[u][j];
                hist[u][j] = (uint32_t)(run - tptr[j]);  // offset within the column
                run += c;
            }
        }
    });
    run_workers(T, [&](size_t t) {
        std::vector<uint32_t> &h = hist[t];
        for (size_t r = cut[t]; r < cut[t + 1]; ++r)
            for (size_t k = rp[r]; k < rp[r + 1]; ++k) {
                const size_t j = ci[k];
                emit(tptr[j] + h[j]++, r, k);
            }
    });
}

void csr_reset(HPRQPMatrixCSR *m)
{
    m->nrow = m->ncol = m->nnz = 0;
    m->row_ptr = nullptr;
    m->col_ind = nullptr;
    m->values = nullptr;
    m->owns_data = 0;
}

double *copy_doubles(const double *src, size_t n)
{
    if (n == 0) return nullptr;
    double *d = static_cast<double *>(std::malloc(n * sizeof(double)));
    if (d != nullptr) std::memcpy(d, src, n * sizeof(double));
    return d;
}

bool offsets_valid(size_t count, const size_t *off, size_t len)
{
    if (count == 0) return true;
    if (count > len || off == nullptr || off[0] >= len || off[count] != len) return false;
    for (size_t k = 0; k < count; ++k)
        if (off[k] >= off[k + 1] || off[k + 1] > len) return false;
    return true;
}

// build.c:51-77
HPRQPStatus validate_cones(const HPRQPModel *md)
{
    if (!offsets_valid(md->num_var_cones, md->var_cone_offsets, md->Q.ncol) ||
        !offsets_valid(md->num_row_cones, md->row_cone_offsets, md->A.nrow))
        return HPRQP_STATUS_INVALID_ARGUMENT;
    if (md->num_var_cones) {
        if (md->l == nullptr || md->u == nullptr) return HPRQP_STATUS_INVALID_ARGUMENT;
        for (size_t i = md->var_cone_offsets[0]; i < md->Q.ncol; ++i)
            if (md->l[i] != -INFINITY || md->u[i] != INFINITY) return HPRQP_STATUS_INVALID_ARGUMENT;
    }
    if (md->num_row_cones) {
        if (md->soc_rhs == nullptr || md->AL == nullptr || md->AU == nullptr) return HPRQP_STATUS_INVALID_ARGUMENT;
        for (size_t i = 0; i < md->A.nrow; ++i) {
            if (!std::isfinite(md->soc_rhs[i])) return HPRQP_STATUS_INVALID_ARGUMENT;
            if (i < md->row_cone_offsets[0]) {
                if (md->soc_rhs[i] != 0.0) return HPRQP_STATUS_INVALID_ARGUMENT;
            } else if (md->AL[i] != -INFINITY || md->AU[i] != INFINITY) {
                return HPRQP_STATUS_INVALID_ARGUMENT;
            }
        }
    }
    return HPRQP_STATUS_OK;
}

HPRQPStatus copy_cone_spec(HPRQPModel *md, const HPRQPSOCSpec *soc)
{
    if (soc == nullptr) return HPRQP_STATUS_OK;
    if (!offsets_valid(soc->num_var_cones, soc->var_offsets, md->Q.ncol) ||
        !offsets_valid(soc->num_row_cones, soc->row_offsets, md->A.nrow) ||
        (soc->num_row_cones && soc->rhs == nullptr))
        return HPRQP_STATUS_INVALID_ARGUMENT;
    if (soc->num_var_cones) {
        size_t bytes = (soc->num_var_cones + 1) * sizeof(size_t);
        md->var_cone_offsets = static_cast<size_t *>(std::malloc(bytes));
        if (!md->var_cone_offsets) return HPRQP_STATUS_ALLOCATION_FAILED;
        std::memcpy(md->var_cone_offsets, soc->var_offsets, bytes);
        md->num_var_cones = soc->num_var_cones;
    }
    if (soc->num_row_cones) {
        size_t bytes = (soc->num_row_cones + 1) * sizeof(size_t);
        md->row_cone_offsets = static_cast<size_t *>(std::malloc(bytes));
        if (!md->row_cone_offsets) return HPRQP_STATUS_ALLOCATION_FAILED;
        std::memcpy(md->row_cone_offsets, soc->row_offsets, bytes);
        md->num_row_cones = soc->num_row_cones;
        md->soc_rhs = copy_doubles(soc->rhs, md->A.nrow);
        if (!md->soc_rhs && md->A.nrow) return HPRQP_STATUS_ALLOCATION_FAILED;
    }
    return validate_cones(md);
}

struct Triple {
    size_t row, col;
    double val;
};

// build.c:157-252: Q_sym = 0.5 (Q + Q'). Every stored entry contributes
// (r, c, v/2) then (c, r, v/2); entries are ordered by (row, col) keeping
// generation order among equal keys, entries with |v| <= 1e-15 are dropped
// before merging, equal keys are summed in order, tiny sums are zeroed and
// zeros removed.
HPRQPStatus symmetrize_q(HPRQPMatrixCSR *out, const HPRQPMatrixCSR *q)
{
    const double tol = 1e-15;
    const size_t n = q->nrow;
    const size_t nnz = q->row_ptr[n];
    // Transposed contributions grouped by target row, each group in
    // generation order (source row ascending, then entry order).
    std::vector<size_t> tptr(n + 1, 0);
    std::vector<size_t> tsrc(nnz);  // source row of each transposed contribution
    std::vector<size_t> tent(nnz);  // its entry index in q
    transpose_scatter(n, n, q->row_ptr, q->col_ind, tptr.data(), [&](size_t dst, size_t r, size_t k) {
        tsrc[dst] = r;
        tent[dst] = k;
    });
    // Row r's contributions in generation order: transposed ones from source
    // rows < r; then row r's own entries, each followed by its own transposed
    // copy when it is diagonal; then transposed ones from source rows > r.
    // A stable sort by column of that sequence is the reference's stable
    // (row, col) order restricted to row r. Row ranges run on worker threads
    // into their own buffers, concatenated in row order.
    const size_t T = host_workers(2 * nnz, (size_t)1 << 19);
    const std::vector<size_t> cut = row_cuts(n, q->row_ptr, T);
    std::vector<size_t> len(n, 0);
    std::vector<std::vector<size_t>> ind_t(T);
    std::vector<std::vector<double>> val_t(T);
    run_workers(T, [&](size_t th) {
        std::vector<size_t> &ind = ind_t[th];
        std::vector<double> &val = val_t[th];
        ind.reserve(2 * (q->row_ptr[cut[th + 1]] - q->row_ptr[cut[th]]));
        val.reserve(ind.capacity());
        std::vector<Triple> row;
        for (size_t r = cut[th]; r < cut[th + 1]; ++r) {
            row.clear();
            size_t t = tptr[r];
            const size_t te = tptr[r + 1];
            for (; t < te && tsrc[t] < r; ++t) row.push_back(Triple{r, tsrc[t], 0.5 * q->values[tent[t]]});
            for (size_t k = q->row_ptr[r]; k < q->row_ptr[r + 1]; ++k) {
                row.push_back(Triple{r, q->col_ind[k], 0.5 * q->values[k]});
                if (q->col_ind[k] == r) {
                    // the transposed copy of this diagonal entry comes right after it
                    row.push_back(Triple{r, r, 0.5 * q->values[tent[t]]});
                    ++t;
                }
            }
            for (; t < te; ++t) row.push_back(Triple{r, tsrc[t], 0.5 * q->values[tent[t]]});
            std::stable_sort(row.begin(), row.end(), [](const Triple &x, const Triple &y) { return x.col < y.col; });
            const size_t row_begin = ind.size();
            for (const Triple &e : row) {
                if (std::fabs(e.val) <= tol) continue;
                if (ind.size() > row_begin && ind.back() == e.col) {
                    val.back() += e.val;
                } else {
                    ind.push_back(e.col);
                    val.push_back(e.val);
                }
            }
            size_t keep = row_begin;
            for (size_t k = row_begin; k < ind.size(); ++k) {
                double v = std::fabs(val[k]) <= tol ? 0.0 : val[k];
                if (v != 0.0) {
                    ind[keep] = ind[k];
                    val[keep] = v;
                    ++keep;
                }
            }
            ind.resize(keep);
            val.resize(keep);
            len[r] = keep - row_begin;
        }
    });
    std::vector<size_t> base(T + 1, 0);
    for (size_t th = 0; th < T; ++th) base[th + 1] = base[th] + ind_t[th].size();
    HPRQPStatus st = matrix_csr_alloc(out, n, q->ncol, base[T]);
    if (st != HPRQP_STATUS_OK) return st;
    for (size_t r = 0; r < n; ++r) out->row_ptr[r + 1] = out->row_ptr[r] + len[r];
    run_workers(T, [&](size_t th) {
        if (ind_t[th].empty()) return;
        std::memcpy(out->col_ind + base[th], ind_t[th].data(), ind_t[th].size() * sizeof(size_t));
        std::memcpy(out->values + base[th], val_t[th].data(), val_t[th].size() * sizeof(double));
    });
    return HPRQP_STATUS_OK;
}

// The transposes and the symmetrisation of a build run on the GPU for large
// models (device_build.h) and on host threads otherwise; both give the same
// bits.
HPRQPStatus build_transpose(HPRQPMatrixCSR *dest, const HPRQPMatrixCSR *src)
{
    if (hprqp::device_transpose_csr(src, dest) == HPRQP_STATUS_OK) return HPRQP_STATUS_OK;
    csr_reset(dest);
    return matrix_csr_transpose(dest, src);
}

HPRQPStatus build_symmetrize(HPRQPMatrixCSR *dest, const HPRQPMatrixCSR *q)
{
    if (hprqp::device_symmetrize_q(q, dest) == HPRQP_STATUS_OK) return HPRQP_STATUS_OK;
    csr_reset(dest);
    return symmetrize_q(dest, q);
}

// build.c:254-337
HPRQPStatus drop_empty_and_free_rows(HPRQPModel *md)
{
    const size_t m = md->A.nrow;
    if (m == 0) return HPRQP_STATUS_OK;
    std::vector<size_t> keep;
    size_t nnz = 0;
    for (size_t r = 0; r < m; ++r) {
        size_t len = md->A.row_ptr[r + 1] - md->A.row_ptr[r];
        bool unbounded = std::isinf(md->AL[r]) && md->AL[r] < 0.0 && std::isinf(md->AU[r]) && md->AU[r] > 0.0;
        if (len > 0 && !unbounded) {
            keep.push_back(r);
            nnz += len;
        }
    }
    if (keep.size() == m) return HPRQP_STATUS_OK;
    HPRQPMatrixCSR na;
    csr_reset(&na);
    HPRQPStatus st = matrix_csr_alloc(&na, keep.size(), md->A.ncol, nnz);
    if (st != HPRQP_STATUS_OK) return st;
    double *nal = keep.empty() ? nullptr : static_cast<double *>(std::malloc(keep.size() * sizeof(double)));
    double *nau = keep.empty() ? nullptr : static_cast<double *>(std::malloc(keep.size() * sizeof(double)));
    if (!keep.empty() && (!nal || !nau)) {
        std::free(nal);
        std::free(nau);
        matrix_csr_free(&na);
        return HPRQP_STATUS_ALLOCATION_FAILED;
    }
    size_t pos = 0;
    for (size_t r = 0; r < keep.size(); ++r) {
        size_t old = keep[r];
        nal[r] = md->AL[old];
        nau[r] = md->AU[old];
        for (size_t k = md->A.row_ptr[old]; k < md->A.row_ptr[old + 1]; ++k) {
            na.col_ind[pos] = md->A.col_ind[k];
            na.values[pos] = md->A.values[k];
            ++pos;
        }
        na.row_ptr[r + 1] = pos;
    }
    matrix_csr_free(&md->A);
    matrix_csr_free(&md->AT);
    std::free(md->AL);
    std::free(md->AU);
    md->A = na;
    md->AL = nal;
    md->AU = nau;
    return build_transpose(&md->AT, &md->A);
}

bool csr_all_finite(const HPRQPMatrixCSR *m)
{
    if (!matrix_csr_is_valid(m)) return false;
    for (size_t k = 0; k < m->nnz; ++k)
        if (!std::isfinite(m->values[k])) return false;
    return true;
}

// build.c:339-400
HPRQPStatus preprocess(HPRQPModel *md)
{
    if (!md->owns_data) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (!matrix_csr_is_valid(&md->Q) || !matrix_csr_is_valid(&md->A) || !matrix_csr_is_valid(&md->AT))
        return HPRQP_STATUS_INVALID_ARGUMENT;
    if (md->Q.nrow != md->Q.ncol || md->Q.nrow != md->A.ncol) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (md->A.nrow > 0 && (md->AL == nullptr || md->AU == nullptr)) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (!csr_all_finite(&md->Q) || !csr_all_finite(&md->A)) return HPRQP_STATUS_INVALID_ARGUMENT;
    for (size_t i = 0; i < md->Q.ncol; ++i)
        if (!std::isfinite(md->c[i])) return HPRQP_STATUS_INVALID_ARGUMENT;
    for (size_t i = 0; i < md->A.nrow; ++i)
        if (std::isnan(md->AL[i]) || std::isnan(md->AU[i])) return HPRQP_STATUS_INVALID_ARGUMENT;
    for (size_t i = 0; i < md->Q.ncol; ++i)
        if (std::isnan(md->l[i]) || std::isnan(md->u[i])) return HPRQP_STATUS_INVALID_ARGUMENT;
    for (size_t i = 0; i < md->Q.ncol; ++i)
        if (md->l[i] > md->u[i]) return HPRQP_STATUS_INVALID_ARGUMENT;
    for (size_t i = 0; i < md->A.nrow; ++i)
        if (md->AL[i] > md->AU[i]) return HPRQP_STATUS_INVALID_ARGUMENT;
    HPRQPStatus st = validate_cones(md);
    if (st != HPRQP_STATUS_OK) return st;
    HPRQPMatrixCSR sym;
    csr_reset(&sym);
    st = build_symmetrize(&sym, &md->Q);
    if (st != HPRQP_STATUS_OK) return st;
    matrix_csr_free(&md->Q);
    md->Q = sym;
    if (md->num_var_cones || md->num_row_cones) return HPRQP_STATUS_OK;
    return drop_empty_and_free_rows(md);
}

// build.c:504-592
HPRQPStatus build_impl(const HPRQPMatrixCSR *Q, const HPRQPMatrixCSR *A, const double *c, const double *AL,
                       const double *AU, const double *l, const double *u, double obj_constant,
                       const HPRQPSOCSpec *soc, HPRQPModel *md)
{
    if (!Q || !A || !md || !c || !l || !u) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (!matrix_csr_is_valid(Q) || !matrix_csr_is_valid(A)) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (Q->nrow != Q->ncol || Q->nrow != A->ncol) return HPRQP_STATUS_INVALID_ARGUMENT;
    const size_t m = A->nrow, n = A->ncol;
    if (m > 0 && (AL == nullptr || AU == nullptr)) return HPRQP_STATUS_INVALID_ARGUMENT;
    model_free(md);
    md->owns_data = 1;
    HPRQPStatus st = matrix_csr_copy(&md->Q, Q);
    if (st == HPRQP_STATUS_OK) st = matrix_csr_copy(&md->A, A);
    if (st == HPRQP_STATUS_OK) st = build_transpose(&md->AT, A);
    if (st == HPRQP_STATUS_OK) {
        md->c = copy_doubles(c, n);
        md->l = copy_doubles(l, n);
        md->u = copy_doubles(u, n);
        md->AL = copy_doubles(AL, m);
        md->AU = copy_doubles(AU, m);
        if ((n && (!md->c || !md->l || !md->u)) || (m && (!md->AL || !md->AU))) st = HPRQP_STATUS_ALLOCATION_FAILED;
    }
    if (st == HPRQP_STATUS_OK) st = copy_cone_spec(md, soc);
    if (st != HPRQP_STATUS_OK) {
        model_free(md);
        return st;
    }
    md->obj_constant = obj_constant;
    md->owns_data = 1;
    return preprocess(md);
}

}  // namespace

extern "C" {

HPRQPStatus matrix_csr_alloc(HPRQPMatrixCSR *matrix, size_t nrow, size_t ncol, size_t nnz)
{
    if (matrix == nullptr) return HPRQP_STATUS_INVALID_ARGUMENT;
    csr_reset(matrix);
    matrix->row_ptr = static_cast<size_t *>(std::calloc(nrow + 1, sizeof(size_t)));
    matrix->col_ind = static_cast<size_t *>(std::calloc(nnz > 0 ? nnz : 1, sizeof(size_t)));
    matrix->values = static_cast<double *>(std::calloc(nnz > 0 ? nnz : 1, sizeof(double)));
    if (!matrix->row_ptr || !matrix->col_ind || !matrix->values) {
        std::free(matrix->row_ptr);
        std::free(matrix->col_ind);
        std::free(matrix->values);
        csr_reset(matrix);
        return HPRQP_STATUS_ALLOCATION_FAILED;
    }
    matrix->nrow = nrow;
    matrix->ncol = ncol;
    matrix->nnz = nnz;
    matrix->owns_data = 1;
    return HPRQP_STATUS_OK;
}

void matrix_csr_free(HPRQPMatrixCSR *matrix)
{
    if (matrix == nullptr) return;
    if (matrix->owns_data) {
        std::free(matrix->row_ptr);
        std::free(matrix->col_ind);
        std::free(matrix->values);
    }
    csr_reset(matrix);
}

int matrix_csr_is_valid(const HPRQPMatrixCSR *matrix)
{
    if (!matrix || !matrix->row_ptr || !matrix->col_ind || !matrix->values) return 0;
    return matrix->row_ptr[0] == 0 && matrix->row_ptr[matrix->nrow] == matrix->nnz;
}

HPRQPStatus matrix_csr_copy(HPRQPMatrixCSR *dest, const HPRQPMatrixCSR *src)
{
    if (!dest || !src || !matrix_csr_is_valid(src)) return HPRQP_STATUS_INVALID_ARGUMENT;
    if (dest == src) return HPRQP_STATUS_OK;
    HPRQPStatus st = matrix_csr_alloc(dest, src->nrow, src->ncol, src->nnz);
    if (st != HPRQP_STATUS_OK) return st;
    parallel_copy(dest->row_ptr, src->row_ptr, (src->nrow + 1) * sizeof(size_t));
    parallel_copy(dest->col_ind, src->col_ind, src->nnz * sizeof(size_t));
    parallel_copy(dest->values, src->values, src->nnz * sizeof(double));
    return HPRQP_STATUS_OK;
}

// sparse.c:81-125: stable counting-sort transpose (rows ascending per column).
HPRQPStatus matrix_csr_transpose(HPRQPMatrixCSR *dest, const HPRQPMatrixCSR *src)
{
    if (!dest || !src || !matrix_csr_is_valid(src)) return HPRQP_STATUS_INVALID_ARGUMENT;
    HPRQPStatus st = matrix_csr_alloc(dest, src->ncol, src->nrow, src->nnz);
    if (st != HPRQP_STATUS_OK) return st;
    transpose_scatter(src->nrow, src->ncol, src->row_ptr, src->col_ind, dest->row_ptr,
                      [&](size_t dst, size_t r, size_t k) {
                          dest->col_ind[dst] = r;
                          dest->values[dst] = src->values[k];
                      });
    return HPRQP_STATUS_OK;
}

// sparse.c:127-145
HPRQPStatus matrix_csr_matvec(const HPRQPMatrixCSR *matrix, const double *x, double *y)
{
    if (!matrix_csr_is_valid(matrix) || !x || !y) return HPRQP_STATUS_INVALID_ARGUMENT;
    for (size_t r = 0; r < matrix->nrow; ++r) {
        double acc = 0.0;
        for (size_t k = matrix->row_ptr[r]; k < matrix->row_ptr[r + 1]; ++k)
            acc += matrix->values[k] * x[matrix->col_ind[k]];
        y[r] = acc;
    }
    return HPRQP_STATUS_OK;
}

// sparse.c:147-167
HPRQPStatus matrix_csr_transpose_matvec(const HPRQPMatrixCSR *matrix, const double *x, double *y)
{
    if (!matrix_csr_is_valid(matrix) || !x || !y) return HPRQP_STATUS_INVALID_ARGUMENT;
    for (size_t j = 0; j < matrix->ncol; ++j) y[j] = 0.0;
    for (size_t r = 0; r < matrix->nrow; ++r) {
        double xr = x[r];
        for (size_t k = matrix->row_ptr[r]; k < matrix->row_ptr[r + 1]; ++k)
            y[matrix->col_ind[k]] += matrix->values[k] * xr;
    }
    return HPRQP_STATUS_OK;
}

void model_init(HPRQPModel *model)
{
    if (model) std::memset(model, 0, sizeof(*model));
}

void model_free(HPRQPModel *model)
{
    if (!model) return;
    if (model->owns_data) {
        matrix_csr_free(&model->Q);
        matrix_csr_free(&model->A);
        matrix_csr_free(&model->AT);
        std::free(model->c);
        std::free(model->AL);
        std::free(model->AU);
        std::free(model->l);
        std::free(model->u);
        std::free(model->var_cone_offsets);
        std::free(model->row_cone_offsets);
        std::free(model->soc_rhs);
    }
    model_init(model);
}

// build.c:434-502
HPRQPStatus model_copy(HPRQPModel *dest, const HPRQPModel *src)
{
    if (!dest || !src || !src->owns_data) return HPRQP_STATUS_INVALID_ARGUMENT;
    model_free(dest);
    dest->owns_data = 1;
    HPRQPStatus st = matrix_csr_copy(&dest->Q, &src->Q);
    if (st == HPRQP_STATUS_OK) st = matrix_csr_copy(&dest->A, &src->A);
    if (st == HPRQP_STATUS_OK) st = matrix_csr_copy(&dest->AT, &src->AT);
    if (st == HPRQP_STATUS_OK) {
        size_t n = src->Q.ncol, m = src->A.nrow;
        dest->c = copy_doubles(src->c, n);
        dest->l = copy_doubles(src->l, n);
        dest->u = copy_doubles(src->u, n);
        dest->AL = copy_doubles(src->AL, m);
        dest->AU = copy_doubles(src->AU, m);
        if ((n && (!dest->c || !dest->l || !dest->u)) || (m && (!dest->AL || !dest->AU)))
            st = HPRQP_STATUS_ALLOCATION_FAILED;
    }
    if (st == HPRQP_STATUS_OK) {
        HPRQPSOCSpec soc = {src->num_var_cones, src->var_cone_offsets, src->num_row_cones, src->row_cone_offsets,
                            src->soc_rhs};
        st = copy_cone_spec(dest, &soc);
    }
    if (st != HPRQP_STATUS_OK) {
        model_free(dest);
        return st;
    }
    dest->obj_constant = src->obj_constant;
    return HPRQP_STATUS_OK;
}

HPRQPStatus build_from_QAbc(const HPRQPMatrixCSR *Q, const HPRQPMatrixCSR *A, const double *c, const double *AL,
                            const double *AU, const double *l, const double *u, double obj_constant,
                            HPRQPModel *model)
{
    return build_impl(Q, A, c, AL, AU, l, u, obj_constant, nullptr, model);
}

HPRQPStatus build_from_QAbc_soc(const HPRQPMatrixCSR *Q, const HPRQPMatrixCSR *A, const double *c,
                                const double *AL, const double *AU, const double *l, const double *u,
                                double obj_constant, const HPRQPSOCSpec *soc, HPRQPModel *model)
{
    return build_impl(Q, A, c, AL, AU, l, u, obj_constant, soc, model);
}

// solver.c:133-157. use_gpu defaults to 1 (this library is the GPU backend).
void params_set_defaults(HPRQPParams *params)
{
    if (!params) return;
    std::memset(params, 0, sizeof(*params));
    params->stoptol = 1e-8;
    params->max_iter = static_cast<size_t>(INT_MAX);
    params->use_gpu = 1;
    params->device_number = 0;
    params->time_limit = 3600.0;
    params->check_iter = 100;
    params->print_frequency = -1;
    params->use_scaling = 1;
    params->verbose = 1;
    params->spmv_mode_q = HPRQP_SPMV_MODE_AUTO;
    params->spmv_mode_a = HPRQP_SPMV_MODE_AUTO;
    params->gram_mode = HPRQP_GRAM_AUTO;
    params->polishing = 1;
    params->polish_max_attempts = 12;
    params->polish_episode_seconds = 20.0;
    params->polish_total_seconds = 0.0;
    params->polish_memory_bytes = static_cast<size_t>(8) * 1024 * 1024 * 1024;
}

void results_init(HPRQPResults *results)
{
    if (results) std::memset(results, 0, sizeof(*results));
}

void results_free(HPRQPResults *results)
{
    if (!results) return;
    std::free(results->x);
    std::free(results->y);
    std::free(results->z);
    std::free(results->w);
    results_init(results);
}

static const char *spmv_mode_name(HPRQPSpmvMode mode)
{
    switch (mode) {
    case HPRQP_SPMV_MODE_AUTO: return "auto";
    case HPRQP_SPMV_MODE_CUSPARSE: return "cusparse";
    case HPRQP_SPMV_MODE_CUSTOMIZED: return "customized";
    default: return "unknown";
    }
}

void print_model_info(const HPRQPModel *model)
{
    if (!model) return;
    std::printf("Model info:\n");
    std::printf("  variables: %zu\n", model->Q.ncol);
    std::printf("  constraints: %zu\n", model->A.nrow);
    std::printf("  nnz(Q): %zu\n", model->Q.nnz);
    std::printf("  nnz(A): %zu\n", model->A.nnz);
}

void print_params_info(const HPRQPParams *params)
{
    if (!params) return;
    std::printf("Parameter info:\n");
    std::printf("  stoptol: %.12g\n", params->stoptol);
    std::printf("  use_gpu: %s\n", params->use_gpu ? "on" : "off");
    std::printf("  device_number: %d\n", params->device_number);
    std::printf("  max_iter: %zu\n", params->max_iter);
    std::printf("  time_limit: %.12g\n", params->time_limit);
    std::printf("  check_iter: %zu\n", params->check_iter);
    std::printf("  print_frequency: %d\n", params->print_frequency);
    std::printf("  use_scaling: %s\n", params->use_scaling ? "on" : "off");
    std::printf("  verbose: %s\n", params->verbose ? "on" : "off");
    std::printf("  spmv mode Q: %s\n", spmv_mode_name(params->spmv_mode_q));
    std::printf("  spmv mode A: %s\n", spmv_mode_name(params->spmv_mode_a));
}

}  // extern "C"
