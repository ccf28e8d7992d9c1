// Times the reference's own spai_gmres_ir (src/ir.cpp:206-227, compiled from
// the reference tree) around the public call, with steady_clock; Matrix
// Market parsing is outside the timed region.  TEST/BENCHMARK INFRASTRUCTURE:
// built by oracle/Makefile into oracle/_ref/ and run by bench.py's CPU
// baseline leg; nothing on the product path runs it.
//
// usage: ref_time MATRIX.mtx EPSILON OUTERCAP INNERCAP WARMUPS TRIALS
//   options as the reference CLI sets them (alpha = ceil(n/8), beta = 8,
//   b_i = 1/sqrt(n)); prints one JSON line: status, outer steps, inner
//   iterations, final normwise backward error and the trial times in seconds.
#include "bspai/cpu.hpp"
#include "bspai/ir.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

int main(int argc, char **argv) {
    if (argc != 7)
        return 2;
    std::ifstream in(argv[1]);
    if (!in)
        return 3;
    const bspai::CsrMatrix a = bspai::read_matrix_market(in);
    bspai::IrOptions o;
    o.spai.epsilon = std::atof(argv[2]);
    o.spai.alpha = a.rows() / 8 + (a.rows() % 8 != 0);
    o.spai.beta = 8;
    o.max_outer_steps = (std::size_t)std::atoll(argv[3]);
    o.gmres.max_iterations = (std::size_t)std::atoll(argv[4]);
    const int warm = std::atoi(argv[5]), trials = std::atoi(argv[6]);
    const std::vector<double> b(a.rows(), 1.0 / std::sqrt(double(a.rows())));
    std::vector<double> times;
    bspai::IrResult last(a.rows());
    for (int t = 0; t < warm + trials; ++t) {
        const auto t0 = std::chrono::steady_clock::now();
        bspai::IrResult r = bspai::spai_gmres_ir(a, b, o);
        const auto t1 = std::chrono::steady_clock::now();
        if (t >= warm)
            times.push_back(std::chrono::duration<double>(t1 - t0).count());
        if (t + 1 == warm + trials)
            last = std::move(r);
    }
    std::printf("{\"status\": \"%s\", \"outer_steps\": %zu, \"inner\": [",
                bspai::ir_status_name(last.status), last.outer_steps);
    for (std::size_t i = 0; i < last.inner_history.size(); ++i)
        std::printf("%s%zu", i ? ", " : "", last.inner_history[i].iterations);
    const double nbe =
        last.history.empty() ? 0.0 : last.history.back().diagnostics.normwise_backward_error;
    std::printf("], \"nbe\": %.17g, \"seconds\": [", nbe);
    for (std::size_t i = 0; i < times.size(); ++i)
        std::printf("%s%.6f", i ? ", " : "", times[i]);
    std::printf("]}\n");
    return 0;
}
