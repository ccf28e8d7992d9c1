// Host CSR container and Matrix Market reader of the HIP backend's C++ API.
// Same names, members and semantics as the reference's include/bspai/cpu.hpp
// (Entry, CsrMatrix::{from_entries, multiply, transpose}, read_matrix_market),
// implemented independently.
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

// Canonical CSR: columns sorted within a row, duplicates summed in input
// order, exact zeros dropped, nonfinite values rejected.
class CsrMatrix {
  public:
    static CsrMatrix from_entries(std::size_t rows, std::size_t cols, std::vector<Entry> entries);
    // Extension (no reference counterpart): wraps arrays that are already
    // canonical, e.g. the output of hip_csr_from_entries; only the sizes are
    // checked.
    static CsrMatrix from_canonical(std::size_t rows, std::size_t cols,
                                    std::vector<std::size_t> offsets,
                                    std::vector<std::size_t> columns, std::vector<double> values);
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

// Matrix Market coordinate real/integer, general/symmetric.
CsrMatrix read_matrix_market(std::istream &input);
// The reader's parse and checks alone: the dimensions and the entries it
// hands to CsrMatrix::from_entries (symmetric files expanded).
std::vector<Entry> parse_matrix_market(std::istream &input, std::size_t &rows, std::size_t &cols);

} // namespace bspai
#endif
