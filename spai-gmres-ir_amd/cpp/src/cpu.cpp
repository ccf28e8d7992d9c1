// Host CSR, Matrix Market reader and status names of the HIP back end's C++
// API.  Semantics follow the reference's src/cpu.cpp:49-161 (canonical CSR,
// strict coordinate parsing); the code is this project's own.
#include "bspai/ir.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>

namespace bspai {

CsrMatrix::CsrMatrix(std::size_t rows, std::size_t cols) : rows_(rows), cols_(cols) {
    if (rows == std::numeric_limits<std::size_t>::max())
        throw std::invalid_argument("Row count overflows CSR offsets");
    offsets_.assign(rows + 1, 0);
}

CsrMatrix CsrMatrix::from_canonical(std::size_t rows, std::size_t cols,
                                    std::vector<std::size_t> offsets,
                                    std::vector<std::size_t> columns, std::vector<double> values) {
    if (offsets.size() != rows + 1 || columns.size() != values.size() ||
        offsets.back() != values.size())
        throw std::invalid_argument("Canonical CSR arrays of inconsistent sizes");
    CsrMatrix m(rows, cols);
    m.offsets_ = std::move(offsets);
    m.columns_ = std::move(columns);
    m.values_ = std::move(values);
    return m;
}

CsrMatrix CsrMatrix::from_entries(std::size_t rows, std::size_t cols, std::vector<Entry> entries) {
    for (const Entry &e : entries) {
        if (e.row >= rows || e.col >= cols)
            throw std::invalid_argument("Entry index outside matrix dimensions");
        if (!std::isfinite(e.value))
            throw std::invalid_argument("Matrix entries must be finite");
    }
    // Order by (row, col); among equal keys keep input order so that
    // duplicates accumulate in the order they were given.
    std::vector<std::size_t> order(entries.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](std::size_t x, std::size_t y) {
        const Entry &a = entries[x], &b = entries[y];
        return a.row != b.row ? a.row < b.row : a.col < b.col;
    });
    CsrMatrix m(rows, cols);
    std::size_t i = 0;
    while (i < order.size()) {
        const Entry &head = entries[order[i]];
        double sum = head.value;
        std::size_t j = i + 1;
        for (; j < order.size() && entries[order[j]].row == head.row &&
               entries[order[j]].col == head.col;
             ++j) {
            sum += entries[order[j]].value;
            if (!std::isfinite(sum))
                throw std::invalid_argument("Duplicate sum is nonfinite");
        }
        if (sum != 0.0) {
            m.columns_.push_back(head.col);
            m.values_.push_back(sum);
            ++m.offsets_[head.row + 1];
        }
        i = j;
    }
    std::partial_sum(m.offsets_.begin(), m.offsets_.end(), m.offsets_.begin());
    return m;
}

std::vector<double> CsrMatrix::multiply(const std::vector<double> &x) const {
    if (x.size() != cols_)
        throw std::invalid_argument("Vector dimension mismatch");
    for (double v : x)
        if (!std::isfinite(v))
            throw std::invalid_argument("Vector must be finite");
    std::vector<double> y(rows_, 0.0);
    for (std::size_t r = 0; r < rows_; ++r) {
        double acc = 0.0;
        for (std::size_t q = offsets_[r]; q < offsets_[r + 1]; ++q)
            acc += values_[q] * x[columns_[q]];
        if (!std::isfinite(acc))
            throw std::overflow_error("Matrix-vector result is nonfinite");
        y[r] = acc;
    }
    return y;
}

CsrMatrix CsrMatrix::transpose() const {
    std::vector<Entry> e;
    e.reserve(values_.size());
    for (std::size_t r = 0; r < rows_; ++r)
        for (std::size_t q = offsets_[r]; q < offsets_[r + 1]; ++q)
            e.push_back(Entry{columns_[q], r, values_[q]});
    return from_entries(cols_, rows_, std::move(e));
}

namespace {
bool data_line(std::istream &in, std::string &line) {
    while (std::getline(in, line)) {
        const std::size_t p = line.find_first_not_of(" \t\r");
        if (p != std::string::npos && line[p] != '%')
            return true;
    }
    return false;
}
std::size_t to_size(const std::string &t) {
    if (t.empty() || t.find_first_not_of("0123456789") != std::string::npos)
        throw std::runtime_error("Expected a nonnegative integer dimension/index");
    std::size_t v = 0;
    for (char ch : t) {
        const std::size_t d = static_cast<std::size_t>(ch - '0');
        if (v > (std::numeric_limits<std::size_t>::max() - d) / 10)
            throw std::runtime_error("Dimension/index exceeds size_t");
        v = v * 10 + d;
    }
    return v;
}
void no_more(std::istringstream &s) {
    std::string extra;
    if (s >> extra)
        throw std::runtime_error("Extra field in Matrix Market line");
}
} // namespace

CsrMatrix read_matrix_market(std::istream &in) {
    std::size_t rows = 0, cols = 0;
    std::vector<Entry> entries = parse_matrix_market(in, rows, cols);
    return CsrMatrix::from_entries(rows, cols, std::move(entries));
}

std::vector<Entry> parse_matrix_market(std::istream &in, std::size_t &rows_out,
                                       std::size_t &cols_out) {
    std::string line, magic, object, format, field, symmetry;
    if (!std::getline(in, line))
        throw std::runtime_error("Missing Matrix Market banner");
    std::istringstream banner(line);
    if (!(banner >> magic >> object >> format >> field >> symmetry))
        throw std::runtime_error("Malformed Matrix Market banner");
    no_more(banner);
    const bool symmetric = symmetry == "symmetric";
    if (magic != "%%MatrixMarket" || object != "matrix" || format != "coordinate" ||
        (field != "real" && field != "integer") || (!symmetric && symmetry != "general"))
        throw std::runtime_error("Unsupported Matrix Market format");
    if (!data_line(in, line))
        throw std::runtime_error("Missing matrix dimensions");
    std::istringstream dims(line);
    std::string tr, tc, tn;
    if (!(dims >> tr >> tc >> tn))
        throw std::runtime_error("Malformed matrix dimensions");
    no_more(dims);
    const std::size_t rows = to_size(tr), cols = to_size(tc), count = to_size(tn);
    if (!rows || !cols)
        throw std::runtime_error("Matrix dimensions must be positive");
    if (symmetric && rows != cols)
        throw std::runtime_error("Symmetric matrix must be square");
    std::vector<Entry> entries;
    entries.reserve(symmetric ? 2 * count : count);
    for (std::size_t i = 0; i < count; ++i) {
        if (!data_line(in, line))
            throw std::runtime_error("Truncated coordinate data");
        std::istringstream s(line);
        std::string ti, tj, tv;
        if (!(s >> ti >> tj >> tv))
            throw std::runtime_error("Malformed coordinate entry");
        no_more(s);
        const std::size_t r = to_size(ti), c = to_size(tj);
        if (!r || !c || r > rows || c > cols)
            throw std::runtime_error("Matrix Market index outside one-based dimensions");
        std::size_t used = 0;
        double v;
        try {
            v = std::stod(tv, &used);
        } catch (const std::exception &) {
            throw std::runtime_error("Invalid coordinate value");
        }
        if (used != tv.size() || !std::isfinite(v))
            throw std::runtime_error("Coordinate value must be finite and numeric");
        if (field == "integer" && tv.find_first_not_of("+-0123456789") != std::string::npos)
            throw std::runtime_error("Integer field requires integer value tokens");
        entries.push_back(Entry{r - 1, c - 1, v});
        if (symmetric && r != c)
            entries.push_back(Entry{c - 1, r - 1, v});
    }
    if (data_line(in, line))
        throw std::runtime_error("Trailing coordinate data");
    rows_out = rows;
    cols_out = cols;
    return entries;
}

const char *spai_status_name(SpaiStatus s) {
    static const char *names[] = {"success",          "not_attempted",    "invalid_input",
                                  "rank_deficient",   "numerical_failure", "empty_candidates",
                                  "zero_denominator", "invalid_score",     "no_growth",
                                  "iteration_limit"};
    return names[static_cast<int>(s)];
}
const char *gmres_status_name(GmresStatus s) {
    static const char *names[] = {"success",
                                  "zero_rhs",
                                  "exact_initial_solution",
                                  "invalid_input",
                                  "zero_preconditioned_rhs",
                                  "breakdown",
                                  "iteration_limit",
                                  "nonconvergence",
                                  "numerical_failure"};
    return names[static_cast<int>(s)];
}
const char *ir_status_name(IrStatus s) {
    static const char *names[] = {"success",    "invalid_input",   "construction_failure",
                                  "inner_failure", "stagnation",  "iteration_limit",
                                  "numerical_failure"};
    return names[static_cast<int>(s)];
}

double ir_backward_tolerance() {
    return 10 * std::ldexp(1.0, -53);
}

SpaiResult::SpaiResult(std::size_t n)
    : scaled_transpose(CsrMatrix::from_entries(n, n, {})),
      inverse(CsrMatrix::from_entries(n, n, {})),
      preconditioner(CsrMatrix::from_entries(n, n, {})) {}

IrResult::IrResult(std::size_t n) : construction(n) {}

} // namespace bspai
