// Iterative-refinement result/option types (names and enum order of the
// reference's include/bspai/ir.hpp:9-62).
#ifndef BSPAI_HIP_CPP_IR_HPP
#define BSPAI_HIP_CPP_IR_HPP

#include "bspai/gmres.hpp"
#include "bspai/spai.hpp"

namespace bspai {

enum class IrStatus {
    success,
    invalid_input,
    construction_failure,
    inner_failure,
    stagnation,
    iteration_limit,
    numerical_failure
};
const char *ir_status_name(IrStatus status);

struct IrOptions {
    SpaiOptions spai;
    GmresOptions gmres;
    std::size_t max_outer_steps = 20;
};

struct IrDiagnostics {
    double residual_inf = 0, residual_2 = 0;
    double normwise_backward_error = 0, componentwise_backward_error = 0;
};

struct IrIteration {
    std::size_t step = 0;
    IrDiagnostics diagnostics;
    bool has_inner = false;
};

struct IrResult {
    explicit IrResult(std::size_t n);
    IrStatus status = IrStatus::invalid_input;
    std::string message;
    bool converged = false, construction_performed = false;
    std::size_t outer_steps = 0, inner_attempts = 0;
    std::vector<double> solution;
    std::vector<IrIteration> history;
    SpaiResult construction;
    std::vector<GmresResult> inner_history;
};

double ir_backward_tolerance(); // 10*2^-53

} // namespace bspai
#endif
