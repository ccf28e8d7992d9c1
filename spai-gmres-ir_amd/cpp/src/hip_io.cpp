// Input helpers of the HIP back end (no counterpart in the reference's
// headers): coordinate entries canonicalized on the GPU, and the Matrix
// Market reader built on the library's GPU parser.  Only the
// reference's public CsrMatrix API is used, so this file compiles against the
// reference's include/bspai too.
#include "bspai/hip_io.hpp"

#include "bspai_hip.h"

#include <climits>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>

namespace bspai {

std::string read_all(std::istream &input) {
    std::string text;
    std::vector<char> buf(1 << 20);
    while (input.read(buf.data(), (std::streamsize)buf.size()) || input.gcount() > 0)
        text.append(buf.data(), (std::size_t)input.gcount());
    return text;
}

CsrMatrix hip_csr_from_entries(std::size_t rows, std::size_t cols,
                               const std::vector<Entry> &entries, int device) {
    bspai_hip_ctx *ctx = nullptr;
    const int rcc = bspai_hip_ctx_create(device, nullptr, &ctx);
    if (rcc != BSPAI_HIP_OK)
        throw std::runtime_error("bspai_hip_ctx_create failed (code " + std::to_string(rcc) + ")");
    std::unique_ptr<bspai_hip_ctx, void (*)(bspai_hip_ctx *)> hold(ctx, bspai_hip_ctx_destroy);
    if (rows >= (std::size_t)INT_MAX || cols >= (std::size_t)INT_MAX ||
        entries.size() >= (std::size_t)INT_MAX)
        throw std::runtime_error("HIP backend: matrix exceeds int32 indexing");
    const std::size_t count = entries.size();
    std::vector<std::int64_t> r(count), cl(count);
    std::vector<double> v(count);
    for (std::size_t i = 0; i < count; ++i) {
        r[i] = (std::int64_t)entries[i].row; // beyond INT64_MAX: negative, out of range
        cl[i] = (std::int64_t)entries[i].col;
        v[i] = entries[i].value;
    }
    std::vector<std::int32_t> ptr(rows + 1), col(count ? count : 1);
    std::vector<double> val(count ? count : 1);
    std::int64_t nnz = 0;
    const int rc = bspai_hip_csr_from_entries(
        ctx, (std::int64_t)rows, (std::int64_t)cols, (std::int64_t)count, r.data(), cl.data(),
        v.data(), BSPAI_HIP_HOST, ptr.data(), col.data(), val.data(), BSPAI_HIP_HOST, &nnz);
    if (rc == BSPAI_HIP_EINVAL)
        throw std::invalid_argument(bspai_hip_last_error(ctx));
    if (rc != BSPAI_HIP_OK)
        throw std::runtime_error(std::string("HIP backend: ") + bspai_hip_last_error(ctx));
    // The canonical entries, row by row: from_entries takes them over in one
    // pass (nothing to sort, sum or drop) and returns the same CSR.
    std::vector<Entry> canon((std::size_t)nnz);
    for (std::size_t k = 0; k < rows; ++k)
        for (std::int32_t q = ptr[k]; q < ptr[k + 1]; ++q)
            canon[(std::size_t)q] = Entry{k, (std::size_t)col[(std::size_t)q], val[(std::size_t)q]};
    return CsrMatrix::from_entries(rows, cols, std::move(canon));
}

CsrMatrix hip_read_matrix_market(std::istream &input, int device) {
    // The text parsed on the GPU (bspai_hip_mm_parse_gpu: the reference
    // reader's rules and messages), the entries canonicalized there too; only
    // the canonical CSR comes back.
    const std::string text = read_all(input);
    bspai_hip_ctx *ctx = nullptr;
    const int rcc = bspai_hip_ctx_create(device, nullptr, &ctx);
    if (rcc != BSPAI_HIP_OK)
        throw std::runtime_error("bspai_hip_ctx_create failed (code " + std::to_string(rcc) + ")");
    std::unique_ptr<bspai_hip_ctx, void (*)(bspai_hip_ctx *)> hold_ctx(ctx, bspai_hip_ctx_destroy);
    bspai_hip_mm_gpu *mm = nullptr;
    char msg[256];
    const int rc = bspai_hip_mm_parse_gpu(ctx, text.data(), (std::int64_t)text.size(), &mm, msg,
                                          (int)sizeof msg);
    if (rc == BSPAI_HIP_EFORMAT)
        throw std::runtime_error(msg);
    if (rc != BSPAI_HIP_OK)
        throw std::runtime_error(std::string("HIP backend: ") + msg);
    std::unique_ptr<bspai_hip_mm_gpu, void (*)(bspai_hip_mm_gpu *)> hold(mm,
                                                                         bspai_hip_mm_gpu_destroy);
    std::int64_t rows = 0, cols = 0, count = 0;
    bspai_hip_mm_gpu_info(mm, &rows, &cols, &count, nullptr, nullptr);
    if (rows >= INT_MAX || cols >= INT_MAX || count >= INT_MAX)
        throw std::runtime_error("HIP backend: matrix exceeds int32 indexing");
    const std::int64_t *r = nullptr, *c = nullptr;
    const double *v = nullptr;
    bspai_hip_mm_gpu_entries(mm, &r, &c, &v);
    std::vector<std::int32_t> ptr((std::size_t)rows + 1), col(count ? (std::size_t)count : 1);
    std::vector<double> val(count ? (std::size_t)count : 1);
    std::int64_t nnz = 0;
    const int rc2 = bspai_hip_csr_from_entries(ctx, rows, cols, count, r, c, v, BSPAI_HIP_DEVICE,
                                               ptr.data(), col.data(), val.data(), BSPAI_HIP_HOST,
                                               &nnz);
    if (rc2 == BSPAI_HIP_EINVAL)
        throw std::invalid_argument(bspai_hip_last_error(ctx));
    if (rc2 != BSPAI_HIP_OK)
        throw std::runtime_error(std::string("HIP backend: ") + bspai_hip_last_error(ctx));
    std::vector<Entry> canon((std::size_t)nnz);
    for (std::size_t k = 0; k < (std::size_t)rows; ++k)
        for (std::int32_t q = ptr[k]; q < ptr[k + 1]; ++q)
            canon[(std::size_t)q] = Entry{k, (std::size_t)col[(std::size_t)q], val[(std::size_t)q]};
    return CsrMatrix::from_entries((std::size_t)rows, (std::size_t)cols, std::move(canon));
}

} // namespace bspai
