// Matrix Market pieces shared by the host reader (mmio.cpp) and the GPU
// parser (mmgpu.hip) inside the library; not part of the C ABI.
#pragma once

#include <cstddef>
#include <stdexcept>

namespace bspai_mmio {

// A malformed file, with the reference reader's message.
struct FormatError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct MmHeader {
    std::size_t rows = 0, cols = 0, count = 0;
    bool integer = false, symmetric = false;
    std::size_t body = 0; // offset of the first line after the dimension line
};

// The banner and the dimension line (reference read_matrix_market,
// src/cpu.cpp:105-130); throws FormatError.
MmHeader mm_read_header(const char *text, std::size_t len);

} // namespace bspai_mmio
