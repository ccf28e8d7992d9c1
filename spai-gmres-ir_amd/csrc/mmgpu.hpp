// Matrix Market text parsed on the GPU (mmgpu.hip).
#pragma once

#include "engine.hpp"

#include <cstdint>

namespace bspai_hip {

struct MmParseShape {
    std::uint64_t rows, cols;
    int integer;
};

// A parsed file: entries in file order (symmetric files expanded), zero-based
// int64 indices and FP64 values, on the device.
struct MmGpu {
    std::int64_t rows = 0, cols = 0, count = 0;
    int integer = 0, symmetric = 0;
    DeviceArray<std::int64_t> row, col;
    DeviceArray<double> val;
};

// Parses text[0, len) (host memory); throws bspai_mmio::FormatError with the
// reference reader's message for a malformed file.
void mm_parse_gpu(Context &c, const char *text, std::size_t len, MmGpu &out);

} // namespace bspai_hip
