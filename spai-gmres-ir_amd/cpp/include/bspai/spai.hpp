// SPAI result/option types (names and enum order of the reference's
// include/bspai/spai.hpp:10-72).
#ifndef BSPAI_HIP_CPP_SPAI_HPP
#define BSPAI_HIP_CPP_SPAI_HPP

#include "bspai/cpu.hpp"

#include <limits>
#include <string>

namespace bspai {

enum class SpaiStatus {
    success,
    not_attempted,
    invalid_input,
    rank_deficient,
    numerical_failure,
    empty_candidates,
    zero_denominator,
    invalid_score,
    no_growth,
    iteration_limit
};
const char *spai_status_name(SpaiStatus status);

struct SpaiOptions {
    double epsilon = 0.3;
    std::size_t alpha = 5;
    std::size_t beta = 8;
};

struct ScoreResult {
    SpaiStatus status = SpaiStatus::not_attempted;
    double score = 0;
    bool clamped = false;
};

// Per-column outcome.  The GPU construction reports the final pattern and
// coefficients and the work counters; per-fit histories are not materialized.
struct SpaiColumn {
    SpaiStatus status = SpaiStatus::not_attempted;
    std::string message;
    std::vector<std::size_t> pattern;
    std::vector<double> coefficients;
    std::size_t fits = 0, growths = 0, score_clamps = 0;
    double reduced_residual = std::numeric_limits<double>::quiet_NaN();
    double full_residual = std::numeric_limits<double>::quiet_NaN();
    double qr_pivot_ratio = std::numeric_limits<double>::quiet_NaN();
};

struct SpaiResult {
    explicit SpaiResult(std::size_t n);
    SpaiStatus status = SpaiStatus::not_attempted;
    std::string message;
    std::vector<double> scaling;
    CsrMatrix scaled_transpose, inverse, preconditioner;
    std::vector<SpaiColumn> columns;
};

} // namespace bspai
#endif
