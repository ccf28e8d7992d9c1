// Input helpers of the HIP back end that have no counterpart in the
// reference's headers: the coordinate-entry canonicalization and the Matrix
// Market reader with the sort and duplicate sums on the GPU.  They build the
// reference's CsrMatrix through its public from_entries, so this header and
// cpp/src/hip_io.cpp compile against the reference's include/bspai as well.
#ifndef BSPAI_HIP_CPP_HIP_IO_HPP
#define BSPAI_HIP_CPP_HIP_IO_HPP

#include "bspai/cpu.hpp"

#include <istream>
#include <string>
#include <vector>

namespace bspai {

// CsrMatrix::from_entries with the sort and the duplicate sums on the GPU
// (bspai_hip_csr_from_entries): the same matrix, and the same
// std::invalid_argument messages for bad entries.
CsrMatrix hip_csr_from_entries(std::size_t rows, std::size_t cols,
                               const std::vector<Entry> &entries, int device = 0);
// read_matrix_market with the text parsed by the library's multi-threaded
// reader (bspai_hip_mm_parse) and the entries canonicalized on the GPU.
CsrMatrix hip_read_matrix_market(std::istream &input, int device = 0);
// The whole remaining content of a stream.
std::string read_all(std::istream &input);

} // namespace bspai
#endif
