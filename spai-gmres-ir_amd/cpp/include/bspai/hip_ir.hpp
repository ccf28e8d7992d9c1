// HIP (gfx950) back end of the SPAI-GMRES-IR solver: the drop-in counterpart
// of the reference's CUDA entry points (include/bspai/cuda_ir.hpp:19-28,
// include/bspai/cuda_spai.hpp:47, include/bspai/cuda_gmres.hpp:40-45).  Every
// call goes through the C ABI of libbspai_hip.so (include/bspai_hip.h); there
// is no CPU numerical fallback.  Runtime/device failures throw
// std::runtime_error, invalid arguments are reported in the result status, as
// in the reference.
#ifndef BSPAI_HIP_CPP_HIP_IR_HPP
#define BSPAI_HIP_CPP_HIP_IR_HPP

#include "bspai/ir.hpp"

namespace bspai {

enum class HipPrecision { fp64, mixed };

struct HipTelemetry {
    double spai_ms = 0, solve_ms = 0;
    std::size_t spai_passes = 0, spai_fits = 0, spai_growths = 0;
    std::size_t p_nnz = 0;
};

struct HipIrResult {
    explicit HipIrResult(std::size_t n) : solver(n) {}
    IrResult solver;
    HipTelemetry telemetry;
};

struct HipSpaiResult {
    explicit HipSpaiResult(std::size_t n) : construction(n) {}
    SpaiResult construction;
    HipTelemetry telemetry;
};

// GPU SPAI construction (reference construct_spai / cuda_construct_spai).
HipSpaiResult hip_construct_spai(const CsrMatrix &a, const SpaiOptions &options,
                                 HipPrecision precision = HipPrecision::fp64, int device = 0);
// One left-preconditioned GMRES solve on the GPU (FP64).
GmresResult hip_left_preconditioned_gmres(const CsrMatrix &a, const CsrMatrix &p,
                                          const std::vector<double> &rhs,
                                          const GmresOptions &options,
                                          const std::vector<double> &initial_guess = {},
                                          int device = 0);
// GPU-resident refinement with a given preconditioner.  The outer loop stops
// on the reference's test (normwise backward error <= 10u, src/ir.cpp:182);
// relres_target > 0 also requires ||b - Ax||_2 <= relres_target ||b||_2.
HipIrResult hip_refine_with_preconditioner(const CsrMatrix &a, const CsrMatrix &p,
                                           const std::vector<double> &rhs,
                                           const IrOptions &options,
                                           HipPrecision precision = HipPrecision::fp64,
                                           int device = 0, double relres_target = 0.0);
// GPU SPAI + GPU-resident IR (reference cuda_gpu_spai_gmres_ir); the stop as
// above.
HipIrResult hip_gpu_spai_gmres_ir(const CsrMatrix &a, const std::vector<double> &rhs,
                                  const IrOptions &options,
                                  HipPrecision precision = HipPrecision::fp64, int device = 0,
                                  double relres_target = 0.0);

} // namespace bspai
#endif
