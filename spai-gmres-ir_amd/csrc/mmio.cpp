// Matrix Market coordinate reader of libbspai_hip.so (host code, several
// threads).  It accepts and rejects the same files as the reference's
// read_matrix_market (src/cpu.cpp:105-161) with the same messages, and hands
// back the coordinate entries in file order (symmetric files expanded), which
// the caller canonicalizes (on the GPU: bspai_hip_csr_from_entries).
//
// Design: the whole text is in memory.  The banner and the dimension line are
// read first; the body is then cut into byte ranges that end at line breaks,
// and one thread per range classifies its lines (blank, comment or data),
// parses its data lines and keeps the first error it meets.  A prefix sum of
// the per-range data-line counts turns local line numbers into global ones:
// the first error among the first `count` data lines is the one a sequential
// reader would have met; without one, fewer data lines than `count` is a
// truncated file and more is trailing data.
//
// Rules kept from the reference (the file contract of the solver CLI):
//   * lines end at '\n'; a line is data unless it is empty after the leading
//     characters " \t\r" or its first other character is '%';
//   * fields are separated by C-locale white space (the >> extraction of a
//     std::istringstream): the banner has five fields, the dimension line and
//     the entries three, and a further field is an error;
//   * indices and dimensions are decimal digit strings that must fit size_t;
//   * a value is what strtod converts, in full, without a range error and
//     finite; an integer field also allows only the characters
//     "+-0123456789".
// The body-line rules and the value conversion are mmtext.hpp's, which the
// GPU parser (mmgpu.hip) shares; mm_read_header is shared with it too.
#include "../../include/bspai_hip.h"
#include "mmio_internal.hpp"
#include "mmtext.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Entry64 {
    std::int64_t row, col;
    double val;
};

using bspai_mmio::FormatError;
using bspai_mmio::MmHeader;

struct Span {
    const char *b = nullptr, *e = nullptr;
    std::size_t size() const { return (std::size_t)(e - b); }
    bool operator==(const char *s) const {
        const std::size_t n = std::strlen(s);
        return size() == n && std::memcmp(b, s, n) == 0;
    }
    bool operator!=(const char *s) const { return !(*this == s); }
};

// Splits [b, e) into at most `want` fields; returns the number found, up to
// want + 1 (one more means "extra field").
int fields(const char *b, const char *e, Span *out, int want) {
    const char *fb[5], *fe[5];
    const int n = bspai_mm::split_fields(b, e, fb, fe, want);
    for (int i = 0; i < n && i < want; ++i)
        out[i] = Span{fb[i], fe[i]};
    return n;
}

// End of the line that starts at p (the position of its '\n', or e).
inline const char *line_end(const char *p, const char *e) {
    const void *nl = std::memchr(p, '\n', (std::size_t)(e - p));
    return nl ? static_cast<const char *>(nl) : e;
}

// A decimal digit string within size_t, else the reference's messages.
std::size_t to_count(const Span &t) {
    std::uint64_t v = 0;
    const int code = bspai_mm::parse_index(t.b, t.e, v);
    if (code)
        throw FormatError(bspai_mm::line_message(code));
    return (std::size_t)v;
}

} // namespace

MmHeader bspai_mmio::mm_read_header(const char *text, std::size_t len) {
    const char *e = text + len;
    MmHeader h;
    if (len == 0)
        throw FormatError("Missing Matrix Market banner");
    const char *le = line_end(text, e);
    Span f[5];
    const int nb = fields(text, le, f, 5);
    if (nb < 5)
        throw FormatError("Malformed Matrix Market banner");
    if (nb > 5)
        throw FormatError("Extra field in Matrix Market line");
    const Span &field = f[3], &sym = f[4];
    if (f[0] != "%%MatrixMarket" || f[1] != "matrix" || f[2] != "coordinate" ||
        (field != "real" && field != "integer") || (sym != "general" && sym != "symmetric"))
        throw FormatError("Unsupported Matrix Market format");
    h.integer = field == "integer";
    h.symmetric = sym == "symmetric";
    const char *p = le < e ? le + 1 : e;
    for (;;) {
        if (p >= e)
            throw FormatError("Missing matrix dimensions");
        const char *q = line_end(p, e);
        if (bspai_mm::is_data_line(p, q)) {
            Span d[3];
            const int nd = fields(p, q, d, 3);
            if (nd < 3)
                throw FormatError("Malformed matrix dimensions");
            if (nd > 3)
                throw FormatError("Extra field in Matrix Market line");
            h.rows = to_count(d[0]);
            h.cols = to_count(d[1]);
            h.count = to_count(d[2]);
            h.body = (std::size_t)((q < e ? q + 1 : e) - text);
            break;
        }
        p = q < e ? q + 1 : e;
    }
    if (h.rows == 0 || h.cols == 0)
        throw FormatError("Matrix dimensions must be positive");
    if (h.symmetric && h.rows != h.cols)
        throw FormatError("Symmetric matrix must be square");
    return h;
}

namespace {

// One byte range of the body, parsed by one thread.
struct Piece {
    const char *b = nullptr, *e = nullptr;
    std::size_t data_lines = 0;
    std::vector<Entry64> entries;
    bool failed = false;
    std::size_t fail_line = 0; // data-line number within the piece
    std::string message;
};

void parse_piece(Piece &pc, const MmHeader &h) {
    const char *p = pc.b;
    while (p < pc.e) {
        const char *q = line_end(p, pc.e);
        if (bspai_mm::is_data_line(p, q)) {
            if (!pc.failed) {
                std::uint64_t r = 0, c = 0;
                double v = 0;
                const int code =
                    bspai_mm::parse_entry_line<true>(p, q, h.rows, h.cols, h.integer, r, c, v);
                if (code) {
                    pc.failed = true;
                    pc.fail_line = pc.data_lines;
                    pc.message = bspai_mm::line_message(code);
                } else {
                    pc.entries.push_back(Entry64{(std::int64_t)r, (std::int64_t)c, v});
                    if (h.symmetric && r != c)
                        pc.entries.push_back(Entry64{(std::int64_t)c, (std::int64_t)r, v});
                }
            }
            ++pc.data_lines;
        }
        p = q < pc.e ? q + 1 : pc.e;
    }
}

} // namespace

struct bspai_hip_mm {
    std::int64_t rows = 0, cols = 0;
    int integer = 0, symmetric = 0;
    std::vector<Entry64> entries;
};

namespace {

void parse_all(const char *text, std::size_t len, int threads, bspai_hip_mm &out) {
    const MmHeader h = bspai_mmio::mm_read_header(text, len);
    const char *b = text + h.body, *e = text + len;
    // Byte ranges of about 4 MiB or more, at most `threads` (16 by default),
    // each ending just after a line break.
    const std::size_t body = (std::size_t)(e - b);
    int t = threads > 0 ? threads : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    t = (int)std::max<std::size_t>(1, std::min<std::size_t>((std::size_t)t, body / (4u << 20) + 1));
    std::vector<Piece> pieces;
    const char *p = b;
    for (int i = 0; i < t && p < e; ++i) {
        const char *q = i + 1 == t ? e : b + body * (std::size_t)(i + 1) / (std::size_t)t;
        if (q < p)
            q = p;
        if (q < e) {
            q = line_end(q, e);
            q = q < e ? q + 1 : e;
        }
        Piece pc;
        pc.b = p;
        pc.e = q;
        pieces.push_back(std::move(pc));
        p = q;
    }
    if (pieces.size() == 1) {
        parse_piece(pieces[0], h);
    } else if (pieces.size() > 1) {
        std::vector<std::thread> pool;
        for (std::size_t i = 1; i < pieces.size(); ++i)
            pool.emplace_back(parse_piece, std::ref(pieces[i]), std::cref(h));
        parse_piece(pieces[0], h);
        for (auto &th : pool)
            th.join();
    }
    // Global order: the first failing data line among the first `count`.
    std::size_t before = 0;
    for (const Piece &pc : pieces) {
        if (pc.failed && before + pc.fail_line < h.count)
            throw FormatError(pc.message);
        before += pc.data_lines;
    }
    if (before < h.count)
        throw FormatError("Truncated coordinate data");
    if (before > h.count)
        throw FormatError("Trailing coordinate data");
    std::size_t total = 0;
    for (const Piece &pc : pieces)
        total += pc.entries.size();
    out.rows = (std::int64_t)h.rows;
    out.cols = (std::int64_t)h.cols;
    out.integer = h.integer;
    out.symmetric = h.symmetric;
    out.entries.clear();
    out.entries.reserve(total);
    for (Piece &pc : pieces) {
        out.entries.insert(out.entries.end(), pc.entries.begin(), pc.entries.end());
        std::vector<Entry64>().swap(pc.entries);
    }
}

void put_message(char *dst, int cap, const char *msg) {
    if (!dst || cap <= 0)
        return;
    std::strncpy(dst, msg, (std::size_t)cap - 1);
    dst[cap - 1] = 0;
}

} // namespace

extern "C" {

int bspai_hip_mm_parse(const char *text, int64_t len, int threads, bspai_hip_mm **out,
                       char *message, int message_len) {
    put_message(message, message_len, "");
    if (!out || len < 0 || (len > 0 && !text)) {
        put_message(message, message_len, "null argument");
        return BSPAI_HIP_EINVAL;
    }
    *out = nullptr;
    try {
        auto mm = new bspai_hip_mm;
        try {
            parse_all(text, (std::size_t)len, threads, *mm);
        } catch (...) {
            delete mm;
            throw;
        }
        *out = mm;
        return BSPAI_HIP_OK;
    } catch (const FormatError &e) {
        put_message(message, message_len, e.what());
        return BSPAI_HIP_EFORMAT;
    } catch (const std::bad_alloc &) {
        put_message(message, message_len, "out of host memory");
        return BSPAI_HIP_ENOMEM;
    } catch (const std::exception &e) {
        put_message(message, message_len, e.what());
        return BSPAI_HIP_EDEVICE;
    }
}

int bspai_hip_mm_info(const bspai_hip_mm *mm, int64_t *rows, int64_t *cols, int64_t *count,
                      int *integer_field, int *symmetric) {
    if (!mm)
        return BSPAI_HIP_EINVAL;
    if (rows)
        *rows = mm->rows;
    if (cols)
        *cols = mm->cols;
    if (count)
        *count = (int64_t)mm->entries.size();
    if (integer_field)
        *integer_field = mm->integer;
    if (symmetric)
        *symmetric = mm->symmetric;
    return BSPAI_HIP_OK;
}

int bspai_hip_mm_entries(const bspai_hip_mm *mm, int64_t *row, int64_t *col, double *val) {
    if (!mm || (!mm->entries.empty() && (!row || !col || !val)))
        return BSPAI_HIP_EINVAL;
    for (std::size_t i = 0; i < mm->entries.size(); ++i) {
        row[i] = mm->entries[i].row;
        col[i] = mm->entries[i].col;
        val[i] = mm->entries[i].val;
    }
    return BSPAI_HIP_OK;
}

void bspai_hip_mm_destroy(bspai_hip_mm *mm) {
    delete mm;
}

} // extern "C"
