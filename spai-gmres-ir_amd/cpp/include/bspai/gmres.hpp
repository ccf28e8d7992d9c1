// GMRES result/option types (names and enum order of the reference's
// include/bspai/gmres.hpp:10-48).
#ifndef BSPAI_HIP_CPP_GMRES_HPP
#define BSPAI_HIP_CPP_GMRES_HPP

#include "bspai/cpu.hpp"

#include <limits>
#include <string>

namespace bspai {

enum class GmresStatus {
    success,
    zero_rhs,
    exact_initial_solution,
    invalid_input,
    zero_preconditioned_rhs,
    breakdown,
    iteration_limit,
    nonconvergence,
    numerical_failure
};
const char *gmres_status_name(GmresStatus status);

struct GmresOptions {
    double tolerance = 1e-8;
    std::size_t max_iterations = 0; // Default n; explicit values must be 1..n.
};

struct GmresResult {
    GmresStatus status = GmresStatus::invalid_input;
    std::string message;
    bool converged = false;
    std::size_t iterations = 0;
    std::vector<double> solution;
    double initial_original_residual = std::numeric_limits<double>::quiet_NaN();
    double initial_preconditioned_relative = std::numeric_limits<double>::quiet_NaN();
    double final_original_residual = std::numeric_limits<double>::quiet_NaN();
    double final_preconditioned_relative = std::numeric_limits<double>::quiet_NaN();
};

} // namespace bspai
#endif
