// Matrix Market dumper linked against the reference's own CPU sources
// (src/cpu.cpp under the reference tree).  TEST INFRASTRUCTURE ONLY: built by
// oracle/Makefile into oracle/_ref/ and run in the build container by
// tests/golden/make_goldens.py and tests/test_mm_reader.py; nothing on the
// product path or on the GPU host runs it.
//
// usage: ref_mm IN.mtx OUT.bin
//   OUT: int64 0, rows, cols, nnz; rows+1 int64 offsets, nnz int64 columns,
//        nnz float64 values -- or int64 kind (1 std::runtime_error,
//        2 std::invalid_argument, 3 other std::exception) then the message.
#include "bspai/cpu.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <stdexcept>

int main(int argc, char **argv) {
    if (argc != 3)
        return 2;
    std::ifstream in(argv[1], std::ios::binary);
    if (!in)
        return 3;
    FILE *out = std::fopen(argv[2], "wb");
    if (!out)
        return 3;
    auto fail = [&](std::int64_t kind, const char *msg) {
        std::fwrite(&kind, 8, 1, out);
        std::fwrite(msg, 1, std::strlen(msg), out);
    };
    try {
        const bspai::CsrMatrix m = bspai::read_matrix_market(in);
        const std::int64_t head[4] = {0, (std::int64_t)m.rows(), (std::int64_t)m.cols(),
                                      (std::int64_t)m.values().size()};
        std::fwrite(head, 8, 4, out);
        for (std::size_t x : m.row_offsets()) {
            const std::int64_t y = (std::int64_t)x;
            std::fwrite(&y, 8, 1, out);
        }
        for (std::size_t x : m.column_indices()) {
            const std::int64_t y = (std::int64_t)x;
            std::fwrite(&y, 8, 1, out);
        }
        std::fwrite(m.values().data(), 8, m.values().size(), out);
    } catch (const std::invalid_argument &e) {
        fail(2, e.what());
    } catch (const std::runtime_error &e) {
        fail(1, e.what());
    } catch (const std::exception &e) {
        fail(3, e.what());
    }
    std::fclose(out);
    return 0;
}
