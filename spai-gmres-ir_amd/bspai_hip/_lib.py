"""ctypes binding of libbspai_hip.so (the C ABI in include/bspai_hip.h).

The library must exist: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libbspai_hip.so")

OK, EINVAL, EDEVICE, ENOMEM, EFORMAT = 0, 1, 2, 3, 4
HOST, DEVICE = 0, 1
FP64, MIXED = 0, 1
SPMV_SLICED, SPMV_SLICED16 = 256, 257  # lanes values of bspai_hip_spmv
MAX_OUTER = 24

SPAI_STATUS = ["success", "not_attempted", "invalid_input", "rank_deficient",
               "numerical_failure", "empty_candidates", "zero_denominator", "invalid_score",
               "no_growth", "iteration_limit"]
GMRES_STATUS = ["success", "zero_rhs", "exact_initial_solution", "invalid_input",
                "zero_preconditioned_rhs", "breakdown", "iteration_limit", "nonconvergence",
                "numerical_failure"]
IR_STATUS = ["success", "invalid_input", "construction_failure", "inner_failure", "stagnation",
             "iteration_limit", "numerical_failure"]


class SpaiOptions(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("alpha", C.c_int64), ("beta", C.c_int64),
                ("precision", C.c_int)]


class SpaiInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("failing_column", C.c_int64), ("n", C.c_int64),
                ("p_nnz", C.c_int64), ("pattern_nnz", C.c_int64), ("fits", C.c_int64),
                ("growths", C.c_int64), ("clamps", C.c_int64), ("passes", C.c_int),
                ("build_ms", C.c_double), ("qr_flops", C.c_double), ("message", C.c_char * 256)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 96), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("bytes", C.c_double), ("flops", C.c_double)]


class GmresOptions(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int64)]


class GmresInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("converged", C.c_int), ("iterations", C.c_int64),
                ("initial_original", C.c_double), ("initial_relative", C.c_double),
                ("final_original", C.c_double), ("final_relative", C.c_double)]


class IrOptions(C.Structure):
    _fields_ = [("spai_epsilon", C.c_double), ("spai_alpha", C.c_int64),
                ("spai_beta", C.c_int64), ("gmres_tolerance", C.c_double),
                ("gmres_max_iterations", C.c_int64), ("max_outer_steps", C.c_int64),
                ("precision", C.c_int), ("relres_target", C.c_double)]


class IrInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("converged", C.c_int),
                ("construction_performed", C.c_int), ("spai_status", C.c_int),
                ("outer_steps", C.c_int64), ("inner_attempts", C.c_int64),
                ("inner_status", C.c_int * MAX_OUTER), ("inner_iterations", C.c_int * MAX_OUTER),
                ("history_len", C.c_int), ("history", (C.c_double * 4) * MAX_OUTER),
                ("residual_inf", C.c_double), ("residual_2", C.c_double),
                ("normwise_backward_error", C.c_double),
                ("componentwise_backward_error", C.c_double), ("spai_ms", C.c_double),
                ("solve_ms", C.c_double), ("message", C.c_char * 256)]


# name -> (restype, argtypes); every symbol of include/bspai_hip.h
_VP = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_F64P = C.c_void_p  # data pointers are passed as raw addresses
SIGNATURES = {
    "bspai_hip_abi_version": (C.c_int, []),
    "bspai_hip_ctx_create": (C.c_int, [C.c_int, _VP, _PP]),
    "bspai_hip_ctx_destroy": (None, [_VP]),
    "bspai_hip_last_error": (C.c_char_p, [_VP]),
    "bspai_hip_synchronize": (C.c_int, [_VP]),
    "bspai_hip_ctx_set_timing": (C.c_int, [_VP, C.c_int]),
    "bspai_hip_ctx_graph_stats": (C.c_int, [_VP, _I64P, _I64P]),
    "bspai_hip_ctx_kernel_times": (C.c_int, [_VP, _VP, C.c_int, C.POINTER(C.c_int)]),
    "bspai_hip_matrix_create": (C.c_int, [_VP, C.c_int64, C.c_int64, _VP, _VP, _VP, C.c_int,
                                          C.c_int, _VP, _PP]),
    "bspai_hip_matrix_destroy": (None, [_VP]),
    "bspai_hip_matrix_info": (C.c_int, [_VP, _I64P, _I64P, C.POINTER(C.c_int)]),
    "bspai_hip_matrix_set_values": (C.c_int, [_VP, _VP, _VP, C.c_int]),
    "bspai_hip_matrix_download": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "bspai_hip_matrix_transpose": (C.c_int, [_VP, _VP, _PP]),
    "bspai_hip_csr_from_entries": (C.c_int, [_VP, C.c_int64, C.c_int64, C.c_int64, _VP, _VP,
                                             _VP, C.c_int, _VP, _VP, _VP, C.c_int,
                                             C.POINTER(C.c_int64)]),
    "bspai_hip_csr_from_entries_noalloc": (C.c_int, [_VP, C.c_int64, C.c_int64, C.c_int64, _VP,
                                                     _VP, _VP, _VP, _VP, _VP,
                                                     C.POINTER(C.c_int64)]),
    "bspai_hip_mm_parse": (C.c_int, [C.c_char_p, C.c_int64, C.c_int, _PP, C.c_char_p, C.c_int]),
    "bspai_hip_mm_info": (C.c_int, [_VP, _I64P, _I64P, _I64P, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]),
    "bspai_hip_mm_entries": (C.c_int, [_VP, _VP, _VP, _VP]),
    "bspai_hip_mm_destroy": (None, [_VP]),
    "bspai_hip_mm_parse_gpu": (C.c_int, [_VP, C.c_char_p, C.c_int64, _PP, C.c_char_p, C.c_int]),
    "bspai_hip_mm_gpu_info": (C.c_int, [_VP, _I64P, _I64P, _I64P, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]),
    "bspai_hip_mm_gpu_entries": (C.c_int, [_VP, _PP, _PP, _PP]),
    "bspai_hip_mm_gpu_copy": (C.c_int, [_VP, _VP, _VP, _VP, _VP, C.c_int]),
    "bspai_hip_mm_gpu_destroy": (None, [_VP]),
    "bspai_hip_spmv": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int, C.c_int]),
    "bspai_hip_spmv_f32": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int, C.c_int]),
    "bspai_hip_spai_build": (C.c_int, [_VP, _VP, C.POINTER(SpaiOptions), _PP]),
    "bspai_hip_spai_destroy": (None, [_VP]),
    "bspai_hip_spai_get_info": (C.c_int, [_VP, C.POINTER(SpaiInfo)]),
    "bspai_hip_spai_preconditioner": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "bspai_hip_spai_columns": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "bspai_hip_spai_patterns": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "bspai_hip_spai_scaling": (C.c_int, [_VP, _VP, _VP]),
    "bspai_hip_score": (C.c_int, [_VP, C.c_double, C.c_double, C.c_double,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]),
    "bspai_hip_gmres": (C.c_int, [_VP, _VP, _VP, _VP, _VP, C.POINTER(GmresOptions), _VP,
                                  C.c_int, C.POINTER(GmresInfo)]),
    "bspai_hip_ir_solve": (C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(IrOptions), _VP, C.c_int,
                                     C.POINTER(IrInfo)]),
    "bspai_hip_spai_gmres_ir": (C.c_int, [_VP, _VP, _VP, C.POINTER(IrOptions), _VP, C.c_int,
                                          C.POINTER(IrInfo)]),
    "bspai_hip_adjoint_solve": (C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(IrOptions), _VP,
                                          C.c_int, C.POINTER(IrInfo)]),
    "bspai_hip_spai_matrix": (C.c_int, [_VP, _VP, C.c_int, _PP]),
    "bspai_hip_sddmm": (C.c_int, [_VP, _VP, C.c_double, _VP, _VP, _VP, C.c_int]),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load libbspai_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is missing: build it with `make -C spai-gmres-ir_amd` "
            "(or __graft_entry__.build()); there is no fallback path")
    # PyTorch for ROCm ships its own HIP and HSA runtimes (sonames
    # libamdhip64.so.7, ...).  Loaded first, the library binds to those, and
    # one runtime serves both; loaded before torch, it would pull in the
    # system runtime, and torch's own second runtime then finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class HipError(RuntimeError):
    pass


def check(rc: int, ctx_ptr) -> None:
    if rc != OK:
        msg = load().bspai_hip_last_error(ctx_ptr)
        kind = {EINVAL: ValueError, ENOMEM: MemoryError}.get(rc, HipError)
        raise kind(msg.decode() if msg else f"bspai_hip error {rc}")
