// Canonical-CSR dumper linked against the reference's own CPU sources
// (src/cpu.cpp under the reference tree).  TEST INFRASTRUCTURE ONLY: built by
// oracle/Makefile into oracle/_ref/ and run in the build container by
// tests/test_oracle.py; nothing on the product path or on the GPU host runs it.
//
// usage: ref_entries IN.bin OUT.bin
//   IN:  int64 rows, cols, count; then count int64 rows, count int64 columns,
//        count float64 values (little-endian).
//   OUT: int64 0, nnz; rows+1 int64 offsets, nnz int64 columns, nnz float64
//        values -- or int64 1 followed by the exception message.
#include "bspai/cpu.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

int main(int argc, char **argv) {
    if (argc != 3)
        return 2;
    FILE *in = std::fopen(argv[1], "rb");
    if (!in)
        return 3;
    std::int64_t head[3];
    if (std::fread(head, sizeof(std::int64_t), 3, in) != 3)
        return 3;
    const std::size_t count = (std::size_t)head[2];
    std::vector<std::int64_t> r(count), c(count);
    std::vector<double> v(count);
    if (std::fread(r.data(), 8, count, in) != count || std::fread(c.data(), 8, count, in) != count ||
        std::fread(v.data(), 8, count, in) != count)
        return 3;
    std::fclose(in);
    std::vector<bspai::Entry> entries(count);
    for (std::size_t i = 0; i < count; ++i)
        entries[i] = bspai::Entry{(std::size_t)r[i], (std::size_t)c[i], v[i]};
    FILE *out = std::fopen(argv[2], "wb");
    if (!out)
        return 3;
    try {
        const bspai::CsrMatrix m =
            bspai::CsrMatrix::from_entries((std::size_t)head[0], (std::size_t)head[1], entries);
        const std::int64_t tag[2] = {0, (std::int64_t)m.values().size()};
        std::fwrite(tag, 8, 2, out);
        for (std::size_t x : m.row_offsets()) {
            const std::int64_t y = (std::int64_t)x;
            std::fwrite(&y, 8, 1, out);
        }
        for (std::size_t x : m.column_indices()) {
            const std::int64_t y = (std::int64_t)x;
            std::fwrite(&y, 8, 1, out);
        }
        std::fwrite(m.values().data(), 8, m.values().size(), out);
    } catch (const std::exception &e) {
        const std::int64_t tag = 1;
        std::fwrite(&tag, 8, 1, out);
        std::fwrite(e.what(), 1, std::strlen(e.what()), out);
    }
    std::fclose(out);
    return 0;
}
