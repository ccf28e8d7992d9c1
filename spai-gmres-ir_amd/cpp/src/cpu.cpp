// Host CSR, Matrix Market reader and status names of the HIP back end's C++
// API.  The contract (which inputs are accepted, the canonical form, the
// exception types and messages) is that of the reference's src/cpu.cpp:43-161;
// the code is this project's own: the reader hands the text to the library's
// multi-threaded parser (bspai_hip_mm_parse), and from_entries recognizes
// input that is already canonical in one pass.
#include "bspai/hip_io.hpp"
#include "bspai/ir.hpp"

#include "bspai_hip.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>

namespace bspai {

CsrMatrix::CsrMatrix(std::size_t rows, std::size_t cols)
    : rows_(rows), cols_(cols), offsets_() {
    // rows + 1 offsets must be countable
    if (rows + 1 == 0)
        throw std::invalid_argument("Row count overflows CSR offsets");
    offsets_.resize(rows + 1, 0);
}

namespace {

bool entry_less(const Entry &a, const Entry &b) {
    return a.row != b.row ? a.row < b.row : a.col < b.col;
}

// Strictly increasing (row, col) and no zero: the canonical form already.
bool canonical(const std::vector<Entry> &e) {
    for (std::size_t i = 0; i < e.size(); ++i) {
        if (e[i].value == 0.0)
            return false;
        if (i && !entry_less(e[i - 1], e[i]))
            return false;
    }
    return true;
}

} // namespace

CsrMatrix CsrMatrix::from_entries(std::size_t rows, std::size_t cols, std::vector<Entry> entries) {
    // The first bad entry in input order decides; its indices before its value.
    const auto bad = std::find_if(entries.begin(), entries.end(), [&](const Entry &e) {
        return e.row >= rows || e.col >= cols || !std::isfinite(e.value);
    });
    if (bad != entries.end())
        throw std::invalid_argument(bad->row >= rows || bad->col >= cols
                                        ? "Entry index outside matrix dimensions"
                                        : "Matrix entries must be finite");
    CsrMatrix m(rows, cols);
    std::vector<std::size_t> &count = m.offsets_;
    if (canonical(entries)) {
        m.columns_.reserve(entries.size());
        m.values_.reserve(entries.size());
        for (const Entry &e : entries) {
            m.columns_.push_back(e.col);
            m.values_.push_back(e.value);
            ++count[e.row + 1];
        }
    } else {
        // A stable order keeps duplicates in input order, which is the order
        // their sum is formed in.
        std::stable_sort(entries.begin(), entries.end(), entry_less);
        auto run = entries.begin();
        while (run != entries.end()) {
            auto next = run + 1;
            double sum = run->value;
            for (; next != entries.end() && next->row == run->row && next->col == run->col;
                 ++next) {
                sum += next->value;
                if (!std::isfinite(sum))
                    throw std::invalid_argument("Duplicate sum is nonfinite");
            }
            if (sum != 0.0) {
                m.columns_.push_back(run->col);
                m.values_.push_back(sum);
                ++count[run->row + 1];
            }
            run = next;
        }
    }
    std::partial_sum(count.begin(), count.end(), count.begin());
    return m;
}

std::vector<double> CsrMatrix::multiply(const std::vector<double> &x) const {
    if (x.size() != cols_)
        throw std::invalid_argument("Vector dimension mismatch");
    if (!std::all_of(x.begin(), x.end(), [](double v) { return std::isfinite(v); }))
        throw std::invalid_argument("Vector must be finite");
    std::vector<double> y(rows_);
    for (std::size_t r = 0; r < rows_; ++r) {
        // the row's products added in stored order, from zero
        double s = 0.0;
        const std::size_t end = offsets_[r + 1];
        for (std::size_t q = offsets_[r]; q != end; ++q)
            s += values_[q] * x[columns_[q]];
        if (!std::isfinite(s))
            throw std::overflow_error("Matrix-vector result is nonfinite");
        y[r] = s;
    }
    return y;
}

CsrMatrix CsrMatrix::transpose() const {
    std::vector<Entry> flipped;
    flipped.reserve(values_.size());
    for (std::size_t r = 0; r < rows_; ++r)
        for (std::size_t q = offsets_[r]; q < offsets_[r + 1]; ++q)
            flipped.push_back(Entry{columns_[q], r, values_[q]});
    return from_entries(cols_, rows_, std::move(flipped));
}

CsrMatrix read_matrix_market(std::istream &input) {
    const std::string text = read_all(input);
    bspai_hip_mm *mm = nullptr;
    char msg[256];
    const int rc = bspai_hip_mm_parse(text.data(), (std::int64_t)text.size(), 0, &mm, msg,
                                      (int)sizeof msg);
    if (rc != BSPAI_HIP_OK)
        throw std::runtime_error(msg);
    std::unique_ptr<bspai_hip_mm, void (*)(bspai_hip_mm *)> hold(mm, bspai_hip_mm_destroy);
    std::int64_t rows = 0, cols = 0, count = 0;
    bspai_hip_mm_info(mm, &rows, &cols, &count, nullptr, nullptr);
    std::vector<std::int64_t> r((std::size_t)count), c((std::size_t)count);
    std::vector<double> v((std::size_t)count);
    bspai_hip_mm_entries(mm, r.data(), c.data(), v.data());
    std::vector<Entry> entries((std::size_t)count);
    for (std::size_t i = 0; i < entries.size(); ++i)
        entries[i] = Entry{(std::size_t)r[i], (std::size_t)c[i], v[i]};
    return CsrMatrix::from_entries((std::size_t)rows, (std::size_t)cols, std::move(entries));
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
