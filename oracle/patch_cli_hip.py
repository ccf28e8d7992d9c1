"""Writes the reference's solver CLI with the HIP flavour a maintainer adds
(the BSPAI_SOLVER_HIP branch beside BSPAI_SOLVER_CUDA that INTEGRATION.md
describes) into oracle/_ref/.  Build-container only: reads the reference tree,
writes nothing under version control.  oracle/Makefile compiles the result
against the reference's own headers and CPU sources plus this repository's
C++ adapter, which is the drop-in the adapter claims to be.

usage: python patch_cli_hip.py REFERENCE/src/solver_main.cpp OUT.cpp
"""
import sys


def patch(text: str) -> str:
    # the adapter's header beside the CUDA one
    inc_old = "#elif defined(BSPAI_SOLVER_CUDA)\n#include \"bspai/cuda_ir.hpp\"\n#endif"
    inc_new = ("#elif defined(BSPAI_SOLVER_HIP)\n#include \"bspai/hip_ir.hpp\"\n"
               "#elif defined(BSPAI_SOLVER_CUDA)\n#include \"bspai/cuda_ir.hpp\"\n#endif")
    # the solve: the HIP entry point returns the same IrResult in .solver
    call_old = ("#ifdef BSPAI_SOLVER_CUDA\n        const auto native = "
                "bspai::cuda_gpu_spai_gmres_ir(matrix, rhs, options);")
    call_new = ("#if defined(BSPAI_SOLVER_HIP)\n        const auto native = "
                "bspai::hip_gpu_spai_gmres_ir(matrix, rhs, options);\n"
                "        const auto &result = native.solver;\n"
                "#elif defined(BSPAI_SOLVER_CUDA)\n        const auto native = "
                "bspai::cuda_gpu_spai_gmres_ir(matrix, rhs, options);")
    # the backend label
    out_old = ("#ifdef BSPAI_SOLVER_CUDA\n                  << \" backend=GPU cpu_outer=true\\n\"")
    out_new = ("#if defined(BSPAI_SOLVER_HIP)\n                  << \" backend=GPU cpu_outer=false\\n\"\n"
               "#elif defined(BSPAI_SOLVER_CUDA)\n                  << \" backend=GPU cpu_outer=true\\n\"")
    for old, new in ((inc_old, inc_new), (call_old, call_new), (out_old, out_new)):
        if text.count(old) != 1:
            raise SystemExit(f"patch_cli_hip: anchor not found once: {old[:60]!r}")
        text = text.replace(old, new)
    return text


if __name__ == "__main__":
    src, dst = sys.argv[1], sys.argv[2]
    with open(src) as f:
        text = f.read()
    with open(dst, "w") as f:
        f.write(patch(text))
