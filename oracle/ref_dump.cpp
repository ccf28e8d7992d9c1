// Golden-vector dumper linked against the reference's own CPU sources
// (src/cpu.cpp, src/spai.cpp, src/gmres.cpp, src/ir.cpp under the reference
// tree).  TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/
// and run in the build container by tests/golden/make_goldens.py; nothing on
// the product path or on the GPU host runs it.
//
// usage: ref_dump MATRIX.mtx RHS(equal|file.bin) eps alpha beta tol maxit outer OUT.json
//   RHS "equal" is the CLI right side b_i = 1/sqrt(n) (src/solver_main.cpp:94);
//   otherwise a raw little-endian float64 file of n values.
// Writes: SPAI status, per-column status/fits/growths/pattern/coefficient bits,
// canonical P (bits), and the full SPAI-GMRES-IR result (src/ir.cpp:206-227).
#include "bspai/ir.hpp"

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {
std::uint64_t bits(double v) {
    std::uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}
void hexs(FILE *f, const std::vector<double> &v) {
    std::fputc('[', f);
    for (std::size_t i = 0; i < v.size(); ++i)
        std::fprintf(f, "%s\"%016" PRIx64 "\"", i ? "," : "", bits(v[i]));
    std::fputc(']', f);
}
void ints(FILE *f, const std::vector<std::size_t> &v) {
    std::fputc('[', f);
    for (std::size_t i = 0; i < v.size(); ++i)
        std::fprintf(f, "%s%zu", i ? "," : "", v[i]);
    std::fputc(']', f);
}
} // namespace

int main(int argc, char **argv) {
    if (argc != 10) {
        std::fprintf(stderr, "usage: %s MATRIX.mtx RHS eps alpha beta tol maxit outer OUT.json\n",
                     argv[0]);
        return 2;
    }
    std::ifstream in(argv[1]);
    const bspai::CsrMatrix a = bspai::read_matrix_market(in);
    const std::size_t n = a.rows();
    std::vector<double> b(n);
    if (std::string(argv[2]) == "equal") {
        for (auto &v : b)
            v = 1.0 / std::sqrt(double(n));
    } else {
        FILE *rf = std::fopen(argv[2], "rb");
        if (!rf || std::fread(b.data(), sizeof(double), n, rf) != n)
            return 3;
        std::fclose(rf);
    }
    bspai::IrOptions o;
    o.spai.epsilon = std::stod(argv[3]);
    o.spai.alpha = std::stoull(argv[4]);
    o.spai.beta = std::stoull(argv[5]);
    o.gmres.tolerance = std::stod(argv[6]);
    o.gmres.max_iterations = std::stoull(argv[7]);
    o.max_outer_steps = std::stoull(argv[8]);
    const bspai::IrResult r = bspai::spai_gmres_ir(a, b, o);
    FILE *f = std::fopen(argv[9], "w");
    const bspai::SpaiResult &s = r.construction;
    std::fprintf(f, "{\"n\":%zu,\"nnz\":%zu,\"spai_status\":\"%s\",\"spai_message\":\"%s\"", n,
                 a.values().size(), bspai::spai_status_name(s.status), s.message.c_str());
    std::fprintf(f, ",\"scaling\":");
    hexs(f, s.scaling);
    std::fprintf(f, ",\"columns\":[");
    for (std::size_t k = 0; k < s.columns.size(); ++k) {
        const auto &c = s.columns[k];
        std::fprintf(f, "%s{\"s\":\"%s\",\"f\":%zu,\"g\":%zu,\"cl\":%zu,\"J\":", k ? "," : "",
                     bspai::spai_status_name(c.status), c.fits, c.growths, c.score_clamps);
        ints(f, c.pattern);
        std::fprintf(f, ",\"c\":");
        hexs(f, c.coefficients);
        std::fprintf(f, ",\"r\":\"%016" PRIx64 "\"}", bits(c.reduced_residual));
    }
    std::fprintf(f, "],\"P\":{\"ptr\":");
    ints(f, s.preconditioner.row_offsets());
    std::fprintf(f, ",\"col\":");
    ints(f, s.preconditioner.column_indices());
    std::fprintf(f, ",\"val\":");
    hexs(f, s.preconditioner.values());
    std::fprintf(f, "},\"ir\":{\"status\":\"%s\",\"converged\":%d,\"outer_steps\":%zu,"
                    "\"inner_attempts\":%zu,\"message\":\"%s\"",
                 bspai::ir_status_name(r.status), r.converged ? 1 : 0, r.outer_steps,
                 r.inner_attempts, r.message.c_str());
    std::fprintf(f, ",\"inner\":[");
    for (std::size_t i = 0; i < r.inner_history.size(); ++i) {
        const auto &g = r.inner_history[i];
        std::fprintf(f, "%s{\"status\":\"%s\",\"iterations\":%zu,\"final_rel\":\"%016" PRIx64
                        "\",\"solution\":",
                     i ? "," : "", bspai::gmres_status_name(g.status), g.iterations,
                     bits(g.final_preconditioned_relative));
        hexs(f, g.solution);
        std::fprintf(f, "}");
    }
    std::fprintf(f, "],\"history\":[");
    for (std::size_t h = 0; h < r.history.size(); ++h) {
        const auto &d = r.history[h].diagnostics;
        std::fprintf(f, "%s[\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64
                        "\"]",
                     h ? "," : "", bits(d.residual_inf), bits(d.residual_2),
                     bits(d.normwise_backward_error), bits(d.componentwise_backward_error));
    }
    std::fprintf(f, "],\"solution\":");
    hexs(f, r.solution);
    std::fprintf(f, "}}\n");
    std::fclose(f);
    return 0;
}
