// Solver command line of the HIP back end:
//   hip_fp64_ir  MATRIX.mtx [epsilon=0.3] [outercap=20] [innercap=0(default n)]
//   hip_mixed_ir (same arguments; FP32 SPAI + FP32 inner GMRES, FP64 outer)
// Argument rules, output lines and exit codes are those of the reference's
// solver CLI (src/solver_main.cpp:60-144): 0 converged, 1 solver failure or
// runtime error, 2 invalid argument / out of range.  The right side is
// b_i = 1/sqrt(n); alpha = ceil(n/8); beta = 8.
#include "bspai/hip_io.hpp"
#include "bspai/hip_ir.hpp"

#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>

#ifdef BSPAI_HIP_CLI_MIXED
static const bspai::HipPrecision kPrecision = bspai::HipPrecision::mixed;
static const char *kPrecisionLabel = "FP32/FP64";
#else
static const bspai::HipPrecision kPrecision = bspai::HipPrecision::fp64;
static const char *kPrecisionLabel = "FP64";
#endif

namespace {
std::size_t parse_cap(const char *text) {
    const std::string s(text);
    if (s.empty() || s.find_first_not_of("0123456789") != std::string::npos)
        throw std::invalid_argument("Expected an unsigned integer cap");
    const unsigned long long v = std::stoull(s);
    if (v > std::numeric_limits<std::size_t>::max())
        throw std::out_of_range("Cap exceeds size_t range");
    return static_cast<std::size_t>(v);
}
} // namespace

int main(int argc, char **argv) {
    if (argc < 2 || argc > 5) {
        std::cerr << "Usage: " << argv[0]
                  << " MATRIX.mtx [epsilon=0.3] [outercap=20] [innercap=0(default n)]\n";
        return 2;
    }
    try {
        std::ifstream in(argv[1]);
        if (!in)
            throw std::invalid_argument("Cannot open Matrix Market input");
        const bspai::CsrMatrix a = bspai::hip_read_matrix_market(in);
        if (!a.rows() || a.rows() != a.cols())
            throw std::invalid_argument("Solver requires a nonempty square matrix");
        bspai::IrOptions o;
        o.spai.alpha = a.rows() / 8 + (a.rows() % 8 != 0);
        o.spai.beta = 8;
        // The mixed flavour runs its correction GMRES in FP32, whose residual
        // stagnates near 1e-6: use the reference's FP32 default tolerance
        // (include/bspai/fp32.hpp:103).
        if (kPrecision == bspai::HipPrecision::mixed)
            o.gmres.tolerance = 1e-4;
        if (argc >= 3) {
            const std::string t(argv[2]);
            std::size_t used = 0;
            const double eps = std::stod(t, &used);
            if (used != t.size() || !std::isfinite(eps) || eps <= 0)
                throw std::invalid_argument("Require a finite positive epsilon in precision range");
            o.spai.epsilon = eps;
        }
        if (argc >= 4)
            o.max_outer_steps = parse_cap(argv[3]);
        if (argc >= 5)
            o.gmres.max_iterations = parse_cap(argv[4]);
        if (o.max_outer_steps > 20 || o.gmres.max_iterations > a.rows())
            throw std::invalid_argument("Require outercap<=20 and innercap<=n");
        const std::vector<double> b(a.rows(), 1.0 / std::sqrt(double(a.rows())));
        const bspai::HipIrResult out = bspai::hip_gpu_spai_gmres_ir(a, b, o, kPrecision);
        const bspai::IrResult &r = out.solver;
        const bspai::IrDiagnostics d =
            r.history.empty() ? bspai::IrDiagnostics() : r.history.back().diagnostics;
        std::cout << std::setprecision(std::numeric_limits<double>::max_digits10)
                  << "precision=" << kPrecisionLabel << " backend=GPU cpu_outer=false\n"
                  << "n=" << a.rows() << " nnz=" << a.values().size()
                  << " epsilon=" << o.spai.epsilon << " tau=" << o.gmres.tolerance << '\n'
                  << "status=" << bspai::ir_status_name(r.status) << " converged=" << r.converged
                  << " outer_steps=" << r.outer_steps << '\n'
                  << "message=" << r.message << '\n';
        if (r.converged)
            std::cout << "residual_inf=" << d.residual_inf
                      << " backward_error=" << d.normwise_backward_error << '\n';
        for (std::size_t i = 0; i < r.solution.size(); ++i)
            std::cout << "x[" << i << "]=" << r.solution[i] << '\n';
        return r.converged ? 0 : 1;
    } catch (const std::invalid_argument &e) {
        std::cerr << "Input error: " << e.what() << '\n';
        return 2;
    } catch (const std::out_of_range &e) {
        std::cerr << "Input error: " << e.what() << '\n';
        return 2;
    } catch (const std::exception &e) {
        std::cerr << "Solver error: " << e.what() << '\n';
        return 1;
    }
}
