// Host CSR container and Matrix Market reader of the HIP back end's C++ API.
// It declares exactly what the reference's include/bspai/cpu.hpp declares
// (Entry, CsrMatrix::{from_entries, rows, cols, row_offsets, column_indices,
// values, multiply, transpose}, read_matrix_market), so code written against
// either header compiles against the other; the implementation
// (cpp/src/cpu.cpp) is this project's own.
#ifndef BSPAI_HIP_CPP_CPU_HPP
#define BSPAI_HIP_CPP_CPU_HPP

#include <cstddef>
#include <istream>
#include <vector>

namespace bspai {

struct Entry {
    std::size_t row;
    std::size_t col;
    double value;
};

// Canonical CSR: columns ascending within a row, duplicates summed in input
// order, exact zeros dropped, nonfinite values rejected.
class CsrMatrix {
  public:
    static CsrMatrix from_entries(std::size_t rows, std::size_t cols, std::vector<Entry> entries);
    std::size_t rows() const { return rows_; }
    std::size_t cols() const { return cols_; }
    const std::vector<std::size_t> &row_offsets() const { return offsets_; }
    const std::vector<std::size_t> &column_indices() const { return columns_; }
    const std::vector<double> &values() const { return values_; }
    std::vector<double> multiply(const std::vector<double> &x) const;
    CsrMatrix transpose() const;

  private:
    CsrMatrix(std::size_t rows, std::size_t cols);
    std::size_t rows_, cols_;
    std::vector<std::size_t> offsets_, columns_;
    std::vector<double> values_;
};

// Matrix Market coordinate real/integer, general/symmetric; malformed input
// throws std::runtime_error with the reference reader's messages.
CsrMatrix read_matrix_market(std::istream &input);

} // namespace bspai
#endif
