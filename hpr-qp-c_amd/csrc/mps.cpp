// build_from_mps: a free-format MPS / QPS reader (sections NAME, OBJSENSE,
// ROWS, COLUMNS, RHS, RANGES, BOUNDS, QUADOBJ, QMATRIX, ENDATA) feeding
// build_from_QAbc. It accepts the files the reference loader accepts
// (src/shared/mps_reader.cpp:453-1283, src/shared/mps_qp.cpp:200-326) and
// builds the same model from them:
//   * the first N row is the objective, later N rows are ignored; row
//     defaults E [0, 0], L (-inf, 0], G [0, inf);
//   * only the first RHS / RANGES / BOUNDS set is used; an RHS on the
//     objective row sets the objective constant to minus its value;
//   * RANGES: E rows widen by the signed value, L rows get [rhs - |R|, rhs],
//     G rows [rhs, rhs + |R|];
//   * bounds LO UP FX FR MI PL (BV LI UI and integer markers make a discrete
//     variable, which is rejected); no bound gives [0, inf), an upper bound
//     alone gives [0, u] or (-inf, u] when u < 0;
//   * QUADOBJ lists one triangle of Q (mirrored here), QMATRIX the full
//     symmetric Q (checked); repeated entries are summed;
//   * QCMATRIX <row> lists x'Hx of a quadratic constraint row; only
//     diagonal H is accepted, on one-sided rows where the constraint is
//     convex (H >= 0 on an L row, H <= 0 on a G row). Such rows leave A and
//     become row cones through build_from_QAbc_qc_diagonal with D = 2H
//     (src/shared/mps_qp.cpp:119-205); rows whose H is all zero stay linear;
//   * maximisation, discrete variables and malformed cards are
//     INVALID_ARGUMENT.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <fstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "hprqp.h"
#include "shared/qcqp.h"

namespace {

constexpr size_t kMaxLine = 1022;  // longest card accepted (the reference reads 1024-byte lines)
constexpr size_t kMaxField = 255;

bool to_number(const std::string &tok, double *out)
{
    std::string t(tok);
    for (char &ch : t)
        if (ch == 'd' || ch == 'D') ch = 'e';
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0' || !std::isfinite(v) || errno == ERANGE) return false;
    *out = v;
    return true;
}

// A vector's data as a non-null pointer (the reference loader allocates its
// arrays even for a model without variables or rows, and build_from_QAbc
// rejects null ones).
const double *data_of(const std::vector<double> &v)
{
    static const double none = 0.0;
    return v.empty() ? &none : v.data();
}

struct Coo {
    std::vector<size_t> r, c;
    std::vector<double> v;
    void add(size_t i, size_t j, double x)
    {
        r.push_back(i);
        c.push_back(j);
        v.push_back(x);
    }
};

// COO -> CSR ordered by (row, column); repeated entries summed in file order.
void coo_to_csr(size_t nrow, size_t ncol, const Coo &coo, std::vector<size_t> &rp, std::vector<size_t> &ci,
                std::vector<double> &val)
{
    const size_t nz = coo.v.size();
    std::vector<size_t> count(ncol + 1, 0), by_col(nz), order(nz);
    for (size_t k = 0; k < nz; ++k) count[coo.c[k] + 1]++;
    for (size_t j = 0; j < ncol; ++j) count[j + 1] += count[j];
    for (size_t k = 0; k < nz; ++k) by_col[count[coo.c[k]]++] = k;  // stable by column
    std::vector<size_t> rcount(nrow + 1, 0);
    for (size_t k = 0; k < nz; ++k) rcount[coo.r[k] + 1]++;
    for (size_t i = 0; i < nrow; ++i) rcount[i + 1] += rcount[i];
    for (size_t k : by_col) order[rcount[coo.r[k]]++] = k;  // stable by row: (row, column, file order)
    rp.assign(nrow + 1, 0);
    ci.clear();
    val.clear();
    ci.reserve(nz + 1);  // never empty: a valid CSR has non-null arrays
    val.reserve(nz + 1);
    size_t prev_r = (size_t)-1, prev_c = (size_t)-1;
    for (size_t k : order) {
        if (coo.r[k] == prev_r && coo.c[k] == prev_c) {
            val.back() += coo.v[k];
            continue;
        }
        ci.push_back(coo.c[k]);
        val.push_back(coo.v[k]);
        rp[coo.r[k] + 1]++;
        prev_r = coo.r[k];
        prev_c = coo.c[k];
    }
    for (size_t i = 0; i < nrow; ++i) rp[i + 1] += rp[i];
}

struct MpsModel {
    // rows: 0 objective, -1 ignored objective, k > 0 constraint k - 1
    std::unordered_map<std::string, long> rows;
    std::unordered_map<std::string, long> vars;
    std::vector<char> row_type;
    std::vector<double> lcon, ucon;
    std::vector<double> c, lvar, uvar;
    std::vector<int> discrete;  // 0 continuous, 1 inside an integer marker, 2 integer bound type
    Coo A, Q;
    struct QcRow {
        size_t row;
        std::vector<std::pair<size_t, double>> diag;  // (column, H_jj) as listed
    };
    std::vector<QcRow> qc;
    std::vector<char> qc_seen;
    double c0 = 0.0;
    bool have_objective = false, maximise = false, full_q = false;
    std::string rhs_set, range_set, bound_set;
};

struct ParseError {
    std::string what;
};

[[noreturn]] void fail(size_t line, const std::string &msg)
{
    throw ParseError{"line " + std::to_string(line) + ": " + msg};
}

long row_of(const MpsModel &M, const std::string &name)
{
    auto it = M.rows.find(name);
    return it == M.rows.end() ? -2 : it->second;
}

void set_rhs(MpsModel &M, size_t line, const std::string &row, double v)
{
    const long r = row_of(M, row);
    if (r == 0) {
        M.c0 = -v;
    } else if (r > 0) {
        const size_t i = (size_t)r - 1;
        if (M.row_type[i] == 'E') M.lcon[i] = M.ucon[i] = v;
        else if (M.row_type[i] == 'L') M.ucon[i] = v;
        else M.lcon[i] = v;
    } else if (r != -1) {
        fail(line, "unknown row " + row);
    }
}

void set_range(MpsModel &M, size_t line, const std::string &row, double v)
{
    const long r = row_of(M, row);
    if (r <= 0) fail(line, r == -2 ? "unknown row " + row : "objective row " + row + " in RANGES");
    const size_t i = (size_t)r - 1;
    if (M.row_type[i] == 'E') {
        if (v >= 0.0) M.ucon[i] += v;
        else M.lcon[i] += v;
    } else if (M.row_type[i] == 'L') {
        M.lcon[i] = M.ucon[i] - std::fabs(v);
    } else {
        M.ucon[i] = M.lcon[i] + std::fabs(v);
    }
}

void add_coefficient(MpsModel &M, size_t line, size_t col, const std::string &row, double v)
{
    const long r = row_of(M, row);
    if (r == 0) M.c[col] += v;
    else if (r > 0) M.A.add((size_t)r - 1, col, v);
    else if (r != -1) fail(line, "unknown row " + row);
}

enum class Section { None, ObjSense, Rows, Columns, Rhs, Ranges, Bounds, Quad, QcMatrix };

void parse(std::istream &in, MpsModel &M)
{
    Section sec = Section::None;
    bool seen_name = false, seen_sense = false, seen_rows = false, seen_cols = false, seen_rhs = false;
    bool seen_bounds = false, seen_ranges = false, seen_quad = false, ended = false, in_integer = false;
    std::string raw;
    size_t line = 0;
    std::vector<std::string> f;
    while (std::getline(in, raw)) {
        ++line;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        if (raw.size() > kMaxLine) fail(line, "line too long");
        if (raw.empty() || raw[0] == '*' || raw[0] == '&') continue;
        const bool header = raw[0] != ' ' && raw[0] != '\t';
        f.clear();
        size_t p = 0;
        while (p < raw.size()) {
            while (p < raw.size() && (raw[p] == ' ' || raw[p] == '\t')) ++p;
            size_t q = p;
            while (q < raw.size() && raw[q] != ' ' && raw[q] != '\t') ++q;
            if (q > p) {
                if (q - p > kMaxField) fail(line, "field too long");
                f.emplace_back(raw.substr(p, q - p));
            }
            p = q;
        }
        if (f.empty()) continue;
        if (header) {
            const std::string &h = f[0];
            const bool named = h == "NAME" || h == "OBJSENSE" || h == "QCMATRIX";
            if (named && f.size() > 2) fail(line, "extra fields after " + h);
            if (h == "NAME") {
                if (seen_name) fail(line, "more than one NAME section");
                seen_name = true;
            } else if (h == "OBJSENSE") {
                if (seen_sense) fail(line, "more than one OBJSENSE section");
                seen_sense = true;
                sec = Section::ObjSense;
                if (f.size() == 2) {
                    if (f[1] == "MAX") M.maximise = true;
                    else if (f[1] != "MIN") fail(line, "unknown objective sense " + f[1]);
                }
            } else if (h == "ROWS") {
                if (seen_rows) fail(line, "more than one ROWS section");
                seen_rows = true;
                sec = Section::Rows;
            } else if (h == "COLUMNS") {
                if (seen_cols) fail(line, "more than one COLUMNS section");
                if (!seen_rows) fail(line, "COLUMNS before ROWS");
                seen_cols = true;
                sec = Section::Columns;
            } else if (h == "RHS") {
                if (seen_rhs) fail(line, "more than one RHS section");
                if (!seen_rows || !seen_cols) fail(line, "RHS before ROWS and COLUMNS");
                seen_rhs = true;
                sec = Section::Rhs;
            } else if (h == "RANGES") {
                if (seen_ranges) fail(line, "more than one RANGES section");
                if (!seen_rows || !seen_cols) fail(line, "RANGES before ROWS and COLUMNS");
                seen_ranges = true;
                sec = Section::Ranges;
            } else if (h == "BOUNDS") {
                if (seen_bounds) fail(line, "more than one BOUNDS section");
                if (!seen_cols) fail(line, "BOUNDS before COLUMNS");
                seen_bounds = true;
                sec = Section::Bounds;
            } else if (h == "QUADOBJ" || h == "QMATRIX") {
                if (seen_quad) fail(line, "more than one QUADOBJ section");
                if (!seen_cols) fail(line, "QUADOBJ before COLUMNS");
                seen_quad = true;
                M.full_q = h == "QMATRIX";
                sec = Section::Quad;
            } else if (h == "QCMATRIX") {
                if (!seen_cols || f.size() != 2) fail(line, "QCMATRIX needs a row name after COLUMNS");
                const long r = row_of(M, f[1]);
                if (r <= 0) fail(line, "QCMATRIX on a non-constraint row " + f[1]);
                if (M.qc_seen.empty()) M.qc_seen.assign(M.row_type.size(), 0);
                if (M.qc_seen[(size_t)r - 1]) fail(line, "second QCMATRIX block for row " + f[1]);
                M.qc_seen[(size_t)r - 1] = 1;
                M.qc.push_back({(size_t)r - 1, {}});
                sec = Section::QcMatrix;
            } else if (h == "ENDATA") {
                ended = true;
                break;
            } else {
                fail(line, "unsupported section " + h);
            }
            continue;
        }
        // data cards
        const size_t nf = f.size();
        if (nf > 6) fail(line, "too many fields");
        double v1 = 0.0, v2 = 0.0;
        switch (sec) {
        case Section::ObjSense:
            if (f[0] == "MAX") M.maximise = true;
            else if (f[0] != "MIN") fail(line, "unknown objective sense " + f[0]);
            break;
        case Section::Rows: {
            if (nf != 2 || f[0].size() != 1 || std::string("NELG").find(f[0][0]) == std::string::npos)
                fail(line, "invalid ROWS card");
            if (f[0] == "N") {
                if (!M.have_objective) {
                    M.have_objective = true;
                    M.rows[f[1]] = 0;
                } else {
                    M.rows[f[1]] = -1;  // further objective rows are ignored
                }
                break;
            }
            if (M.rows.count(f[1])) fail(line, "duplicate row " + f[1]);
            M.rows[f[1]] = (long)M.row_type.size() + 1;
            const char t = f[0][0];
            M.row_type.push_back(t);
            M.lcon.push_back(t == 'L' ? -INFINITY : 0.0);
            M.ucon.push_back(t == 'G' ? INFINITY : 0.0);
            break;
        }
        case Section::Columns: {
            if (nf >= 2 && f[1] == "'MARKER'") {
                if (nf >= 3 && f[2] == "'INTORG'") in_integer = true;
                else if (nf >= 3 && f[2] == "'INTEND'") in_integer = false;
                else fail(line, "unknown marker");
                break;
            }
            if ((nf != 3 && nf != 5) || !to_number(f[2], &v1) || (nf == 5 && !to_number(f[4], &v2)))
                fail(line, "invalid coefficient card");
            auto it = M.vars.find(f[0]);
            size_t col;
            if (it == M.vars.end()) {
                col = M.c.size();
                M.vars.emplace(f[0], (long)col);
                M.c.push_back(0.0);
                M.lvar.push_back(NAN);
                M.uvar.push_back(NAN);
                M.discrete.push_back(in_integer ? 1 : 0);
            } else {
                col = (size_t)it->second;
            }
            add_coefficient(M, line, col, f[1], v1);
            if (nf == 5) add_coefficient(M, line, col, f[3], v2);
            break;
        }
        case Section::Rhs:
        case Section::Ranges: {
            if ((nf != 3 && nf != 5) || !to_number(f[2], &v1) || (nf == 5 && !to_number(f[4], &v2)))
                fail(line, "invalid coefficient card");
            std::string &set = sec == Section::Rhs ? M.rhs_set : M.range_set;
            if (set.empty()) set = f[0];
            else if (set != f[0]) break;  // only the first set is used
            auto apply = sec == Section::Rhs ? set_rhs : set_range;
            apply(M, line, f[1], v1);
            if (nf == 5) apply(M, line, f[3], v2);
            break;
        }
        case Section::Bounds: {
            if (nf < 3 || nf > 4 || (nf == 4 && !to_number(f[3], &v1))) fail(line, "invalid bound card");
            if (M.bound_set.empty()) M.bound_set = f[1];
            else if (M.bound_set != f[1]) break;
            auto it = M.vars.find(f[2]);
            if (it == M.vars.end()) fail(line, "unknown column " + f[2]);
            const size_t j = (size_t)it->second;
            const std::string &t = f[0];
            if (t == "FR") {
                M.lvar[j] = -INFINITY;
                M.uvar[j] = INFINITY;
            } else if (t == "MI") {
                M.lvar[j] = -INFINITY;
            } else if (t == "PL") {
                M.uvar[j] = INFINITY;
            } else if (t == "BV") {
                M.discrete[j] = 2;
                M.lvar[j] = 0.0;
                M.uvar[j] = 1.0;
            } else {
                if (nf < 4) fail(line, t + " bound needs a value");
                if (t == "LO") M.lvar[j] = v1;
                else if (t == "UP") M.uvar[j] = v1;
                else if (t == "FX") M.lvar[j] = M.uvar[j] = v1;
                else if (t == "LI") M.discrete[j] = 2, M.lvar[j] = v1;
                else if (t == "UI") M.discrete[j] = 2, M.uvar[j] = v1;
                else fail(line, "unknown bound type " + t);
            }
            break;
        }
        case Section::Quad: {
            if (nf != 3 || !to_number(f[2], &v1)) fail(line, "invalid coefficient card");
            auto a = M.vars.find(f[0]), b = M.vars.find(f[1]);
            if (a == M.vars.end() || b == M.vars.end()) fail(line, "unknown variable in QUADOBJ");
            M.Q.add((size_t)a->second, (size_t)b->second, v1);
            break;
        }
        case Section::QcMatrix: {
            if (nf != 3 || !to_number(f[2], &v1)) fail(line, "invalid QCMATRIX coefficient");
            auto a = M.vars.find(f[0]), b = M.vars.find(f[1]);
            if (a == M.vars.end() || b == M.vars.end()) fail(line, "unknown variable in QCMATRIX");
            if (a->second != b->second && v1 != 0.0) fail(line, "off-diagonal QCMATRIX entry");
            if (v1 != 0.0) M.qc.back().diag.emplace_back((size_t)a->second, v1);
            break;
        }
        default:
            fail(line, "data outside an MPS section");
        }
    }
    if (!ended) fail(line, "end of file before ENDATA");
    // default bounds
    for (size_t j = 0; j < M.c.size(); ++j) {
        const bool nl = std::isnan(M.lvar[j]), nu = std::isnan(M.uvar[j]);
        if (nl && nu) {
            M.lvar[j] = 0.0;
            M.uvar[j] = M.discrete[j] == 1 ? 1.0 : INFINITY;
        } else if (nl) {
            M.lvar[j] = M.uvar[j] < 0.0 ? -INFINITY : 0.0;
        } else if (nu) {
            M.uvar[j] = INFINITY;
        }
    }
}

// Quadratic constraint rows leave A and become the diagonal lift's C, D and
// constants (src/shared/mps_qp.cpp:123-205). A row's H entries are summed
// per column in ascending (column, value) order; zero sums are dropped, and a
// row left without entries stays linear. D = 2H times the row's sense sign
// (+1 for L rows, -1 for G rows), which must come out non-negative.
HPRQPStatus build_quadratic_rows(MpsModel &M, const HPRQPMatrixCSR &Q, const HPRQPMatrixCSR &A, HPRQPModel *model)
{
    const size_t m = A.nrow, n = A.ncol;
    std::vector<const MpsModel::QcRow *> of_row(m, nullptr);
    for (auto &q : M.qc) {
        auto &d = q.diag;
        std::sort(d.begin(), d.end());
        size_t w = 0;
        for (const auto &e : d) {
            if (w && d[w - 1].first == e.first) d[w - 1].second += e.second;
            else d[w++] = e;
        }
        d.resize(w);
        d.erase(std::remove_if(d.begin(), d.end(), [](const auto &e) { return e.second == 0.0; }), d.end());
        if (!d.empty()) of_row[q.row] = &q;
    }
    std::vector<size_t> lrp{0}, lci, crp{0}, cci, drp{0}, dci;
    std::vector<double> lval, cval, dval, lo, hi, constant;
    for (auto *v : {&lci, &cci, &dci}) v->reserve(A.nnz + 1);  // never empty
    for (auto *v : {&lval, &cval, &dval}) v->reserve(A.nnz + 1);
    for (size_t r = 0; r < m; ++r) {
        if (!of_row[r]) {
            for (size_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k) {
                lci.push_back(A.col_ind[k]);
                lval.push_back(A.values[k]);
            }
            lrp.push_back(lci.size());
            lo.push_back(M.lcon[r]);
            hi.push_back(M.ucon[r]);
            continue;
        }
        double sign;
        if (M.lcon[r] == -INFINITY && std::isfinite(M.ucon[r])) sign = 1.0;
        else if (M.ucon[r] == INFINITY && std::isfinite(M.lcon[r])) sign = -1.0;
        else {
            std::fprintf(stderr, "Error: a QCMATRIX row needs a one-sided inequality (no equality, range or free row).\n");
            return HPRQP_STATUS_INVALID_ARGUMENT;
        }
        for (const auto &e : of_row[r]->diag) {
            const double dv = e.second * (2.0 * sign);
            if (!std::isfinite(dv) || dv < 0.0) {
                std::fprintf(stderr, "Error: a QCMATRIX row is not convex in its inequality sense.\n");
                return HPRQP_STATUS_INVALID_ARGUMENT;
            }
            dci.push_back(e.first);
            dval.push_back(dv);
        }
        drp.push_back(dci.size());
        for (size_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k) {
            cci.push_back(A.col_ind[k]);
            cval.push_back(sign * A.values[k]);
        }
        crp.push_back(cci.size());
        constant.push_back(sign > 0.0 ? -M.ucon[r] : M.lcon[r]);
    }
    const size_t nq = constant.size(), nl = lo.size();
    const HPRQPMatrixCSR L{nl, n, lci.size(), lrp.data(), lci.data(), lval.data(), 0};
    const HPRQPMatrixCSR Cm{nq, n, cci.size(), crp.data(), cci.data(), cval.data(), 0};
    const HPRQPMatrixCSR Dm{nq, n, dci.size(), drp.data(), dci.data(), dval.data(), 0};
    return build_from_QAbc_qc_diagonal(&Q, &L, data_of(M.c), data_of(lo), data_of(hi), data_of(M.lvar),
                                       data_of(M.uvar), M.c0, &Cm, data_of(constant), &Dm, model);
}

}  // namespace

extern "C" HPRQPStatus build_from_mps(const char *mps_path, HPRQPModel *model)
{
    if (mps_path == nullptr || model == nullptr) return HPRQP_STATUS_INVALID_ARGUMENT;
    model_init(model);
    std::ifstream in(mps_path);
    if (!in) {
        std::fprintf(stderr, "Error: cannot open %s\n", mps_path);
        return HPRQP_STATUS_INVALID_ARGUMENT;
    }
    MpsModel M;
    try {
        parse(in, M);
    } catch (const ParseError &e) {
        std::fprintf(stderr, "Error: %s: %s\n", mps_path, e.what.c_str());
        return HPRQP_STATUS_INVALID_ARGUMENT;
    } catch (const std::bad_alloc &) {
        return HPRQP_STATUS_ALLOCATION_FAILED;
    }
    if (M.maximise) {
        std::fprintf(stderr, "Error: maximisation MPS input is unsupported; supply a convex minimisation model.\n");
        return HPRQP_STATUS_INVALID_ARGUMENT;
    }
    for (int d : M.discrete)
        if (d) {
            std::fprintf(stderr, "Error: discrete MPS variables are unsupported.\n");
            return HPRQP_STATUS_INVALID_ARGUMENT;
        }
    const size_t n = M.c.size(), m = M.row_type.size();
    try {
        std::vector<size_t> arp, aci, qrp, qci;
        std::vector<double> aval, qval;
        coo_to_csr(m, n, M.A, arp, aci, aval);
        coo_to_csr(n, n, M.Q, qrp, qci, qval);
        if (M.full_q) {
            // QMATRIX holds the whole symmetric matrix: check it
            for (size_t r = 0; r < n; ++r)
                for (size_t k = qrp[r]; k < qrp[r + 1]; ++k) {
                    const size_t cc = qci[k];
                    double mirror = 0.0;
                    for (size_t t = qrp[cc]; t < qrp[cc + 1]; ++t)
                        if (qci[t] == r) mirror = qval[t];
                    if (mirror != qval[k]) {
                        std::fprintf(stderr, "Error: QMATRIX must contain a symmetric full objective matrix.\n");
                        return HPRQP_STATUS_INVALID_ARGUMENT;
                    }
                }
        } else {
            // QUADOBJ holds one triangle: mirror the off-diagonal entries
            Coo full;
            for (size_t r = 0; r < n; ++r)
                for (size_t k = qrp[r]; k < qrp[r + 1]; ++k) {
                    full.add(r, qci[k], qval[k]);
                    if (qci[k] != r) full.add(qci[k], r, qval[k]);
                }
            coo_to_csr(n, n, full, qrp, qci, qval);
        }
        HPRQPMatrixCSR Qm{n, n, qci.size(), qrp.data(), qci.data(), qval.data(), 0};
        HPRQPMatrixCSR Am{m, n, aci.size(), arp.data(), aci.data(), aval.data(), 0};
        if (!M.qc.empty()) return build_quadratic_rows(M, Qm, Am, model);
        return build_from_QAbc(&Qm, &Am, data_of(M.c), data_of(M.lcon), data_of(M.ucon), data_of(M.lvar), data_of(M.uvar),
                               M.c0, model);
    } catch (const std::bad_alloc &) {
        return HPRQP_STATUS_ALLOCATION_FAILED;
    }
}
