"""ctypes front end of the CPU FP64 oracle (liboracle.so).  TEST INFRASTRUCTURE.

Imported only by tests/, by the smoke check and by the benchmark's CPU-baseline
leg, always as the checker -- never on the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SPAI_STATUS = ["success", "not_attempted", "invalid_input", "rank_deficient",
               "numerical_failure", "empty_candidates", "zero_denominator",
               "invalid_score", "no_growth", "iteration_limit"]
GMRES_STATUS = ["success", "zero_rhs", "exact_initial_solution", "invalid_input",
                "zero_preconditioned_rhs", "breakdown", "iteration_limit",
                "nonconvergence", "numerical_failure"]
IR_STATUS = ["success", "invalid_input", "construction_failure", "inner_failure",
             "stagnation", "iteration_limit", "numerical_failure"]

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i64p = C.POINTER(C.c_int64)
        i32p = C.POINTER(C.c_int32)
        f64p = C.POINTER(C.c_double)
        L.oracle_spai_build.restype = P
        L.oracle_spai_build.argtypes = [C.c_int64, i64p, i64p, f64p, C.c_double, C.c_int64,
                                        C.c_int64, C.c_int]
        L.oracle_spai_free.argtypes = [P]
        L.oracle_spai_status.argtypes = [P]
        L.oracle_spai_message.argtypes = [P]
        L.oracle_spai_message.restype = C.c_char_p
        L.oracle_spai_p_nnz.argtypes = [P]
        L.oracle_spai_p_nnz.restype = C.c_int64
        L.oracle_spai_p.argtypes = [P, i64p, i64p, f64p]
        L.oracle_spai_scaling.argtypes = [P, f64p]
        L.oracle_spai_columns.argtypes = [P, i32p, i32p, i32p, i32p, f64p, f64p]
        L.oracle_spai_pattern_nnz.argtypes = [P]
        L.oracle_spai_pattern_nnz.restype = C.c_int64
        L.oracle_spai_patterns.argtypes = [P, i64p, i64p, f64p]
        L.oracle_spai_trace.argtypes = [P, C.c_int64, C.c_int64, C.c_int, i64p]
        L.oracle_spai_trace.restype = C.c_int64
        L.oracle_spai_trace_scores.argtypes = [P, C.c_int64, C.c_int64, f64p]
        L.oracle_spai_trace_scores.restype = C.c_int64
        L.oracle_score.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]
        L.oracle_score.restype = C.c_double
        L.oracle_hypot_chain.argtypes = [f64p, C.c_int64]
        L.oracle_hypot_chain.restype = C.c_double
        L.oracle_gmres.argtypes = [C.c_int64, i64p, i64p, f64p, i64p, i64p, f64p, f64p, f64p,
                                   C.c_double, C.c_int64, f64p, i64p, f64p]
        L.oracle_ir.restype = P
        L.oracle_ir.argtypes = [C.c_int64, i64p, i64p, f64p, i64p, i64p, f64p, f64p, C.c_double,
                                C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_int64,
                                C.c_double]
        L.oracle_ir_free.argtypes = [P]
        L.oracle_ir_info.argtypes = [P, i64p]
        L.oracle_ir_message.argtypes = [P]
        L.oracle_ir_message.restype = C.c_char_p
        L.oracle_ir_solution.argtypes = [P, f64p]
        L.oracle_ir_history.argtypes = [P, f64p]
        L.oracle_ir_inner.argtypes = [P, i64p]
        L.oracle_ir_p_nnz.argtypes = [P]
        L.oracle_ir_p_nnz.restype = C.c_int64
        L.oracle_ir_p.argtypes = [P, i64p, i64p, f64p]
        L.oracle_spai_subset.restype = C.c_int64
        L.oracle_spai_subset.argtypes = [C.c_int64, i64p, i64p, f64p, C.c_double, C.c_int64,
                                         C.c_int64, C.c_int64, i64p, i32p, i32p, i32p, i64p, i64p,
                                         f64p, C.c_int64]
        L.oracle_from_entries.restype = P
        L.oracle_from_entries.argtypes = [C.c_int64, C.c_int64, C.c_int64, i64p, i64p, f64p]
        L.oracle_csr_nnz.argtypes = [P]
        L.oracle_csr_nnz.restype = C.c_int64
        L.oracle_csr_get.argtypes = [P, i64p, i64p, f64p]
        L.oracle_csr_free.argtypes = [P]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_set_progress.argtypes = [C.c_void_p]
        _lib = L
    return _lib


_PROGRESS = C.CFUNCTYPE(None, C.c_int, C.c_int64, C.c_int64)
_progress_keep = None


def set_progress(fn) -> None:
    """fn(phase, done, value) after every 2^19 SPAI columns (phase 0, value =
    n) and every refinement step (phase 1, value = its inner iterations);
    None turns it off."""
    global _progress_keep
    _progress_keep = None if fn is None else _PROGRESS(fn)
    lib().oracle_set_progress(C.cast(_progress_keep, C.c_void_p) if fn else None)


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _csr_args(m):
    ptr = np.ascontiguousarray(m.ptr, dtype=np.int64)
    col = np.ascontiguousarray(m.col, dtype=np.int64)
    val = np.ascontiguousarray(m.val, dtype=np.float64)
    return (ptr, col, val), (_p(ptr, C.c_int64), _p(col, C.c_int64), _p(val, C.c_double))


@dataclass
class SpaiResult:
    status: str
    message: str
    scaling: np.ndarray
    p_ptr: np.ndarray
    p_col: np.ndarray
    p_val: np.ndarray
    col_status: np.ndarray
    fits: np.ndarray
    growths: np.ndarray
    clamps: np.ndarray
    residual: np.ndarray
    pivot: np.ndarray
    pat_ptr: np.ndarray
    pat_idx: np.ndarray
    pat_coef: np.ndarray
    handle: object = field(default=None, repr=False)

    def trace(self, k: int, f: int, which: str):
        L = lib()
        w = {"shadow": 0, "candidates": 1, "selected": 2}[which]
        size = L.oracle_spai_trace(self.handle, k, f, w, None)
        if size < 0:
            return None
        out = np.empty(size, dtype=np.int64)
        L.oracle_spai_trace(self.handle, k, f, w, _p(out, C.c_int64))
        return out

    def trace_scores(self, k: int, f: int):
        L = lib()
        size = L.oracle_spai_trace_scores(self.handle, k, f, None)
        if size < 0:
            return None
        out = np.empty(size)
        L.oracle_spai_trace_scores(self.handle, k, f, _p(out, C.c_double))
        return out

    def __del__(self):
        if self.handle is not None and _lib is not None:
            _lib.oracle_spai_free(self.handle)
            self.handle = None


def construct_spai(a, epsilon=0.3, alpha=None, beta=8, trace=False) -> SpaiResult:
    """reference construct_spai (src/spai.cpp:314-374); alpha defaults to ceil(n/8)."""
    L = lib()
    n = a.n
    if alpha is None:
        alpha = -(-n // 8)
    keep, (pp, pc, pv) = _csr_args(a)
    h = L.oracle_spai_build(n, pp, pc, pv, float(epsilon), int(alpha), int(beta), int(trace))
    if not h:
        raise RuntimeError(L.oracle_last_error().decode())
    pn = L.oracle_spai_p_nnz(h)
    p_ptr = np.empty(n + 1, np.int64)
    p_col = np.empty(pn, np.int64)
    p_val = np.empty(pn)
    L.oracle_spai_p(h, _p(p_ptr, C.c_int64), _p(p_col, C.c_int64), _p(p_val, C.c_double))
    sc = np.empty(n)
    st = L.oracle_spai_status(h)
    if SPAI_STATUS[st] != "invalid_input" or True:
        try:
            L.oracle_spai_scaling(h, _p(sc, C.c_double))
        except Exception:  # pragma: no cover
            pass
    cs = np.zeros(n, np.int32)
    fi = np.zeros(n, np.int32)
    gr = np.zeros(n, np.int32)
    cl = np.zeros(n, np.int32)
    rs = np.full(n, np.nan)
    pv_ = np.full(n, np.nan)
    pat_ptr = np.zeros(n + 1, np.int64)
    pat_idx = np.zeros(0, np.int64)
    pat_coef = np.zeros(0)
    if L.oracle_spai_status(h) != 2:  # columns exist unless input was invalid
        L.oracle_spai_columns(h, _p(cs, C.c_int32), _p(fi, C.c_int32), _p(gr, C.c_int32),
                              _p(cl, C.c_int32), _p(rs, C.c_double), _p(pv_, C.c_double))
        tn = L.oracle_spai_pattern_nnz(h)
        pat_idx = np.empty(tn, np.int64)
        pat_coef = np.empty(tn)
        L.oracle_spai_patterns(h, _p(pat_ptr, C.c_int64), _p(pat_idx, C.c_int64),
                               _p(pat_coef, C.c_double))
    res = SpaiResult(SPAI_STATUS[st], L.oracle_spai_message(h).decode(), sc, p_ptr, p_col,
                     p_val, cs, fi, gr, cl, rs, pv_, pat_ptr, pat_idx, pat_coef)
    res.handle = h
    return res


def spai_subset(a, cols, epsilon=0.3, alpha=None, beta=8):
    """construct_column (src/spai.cpp:100-244) for the listed columns only.

    Returns (status[m], fits[m], growths[m], ptr[m+1], idx, coef)."""
    L = lib()
    n = a.n
    if alpha is None:
        alpha = -(-n // 8)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    m = cols.size
    keep, (pp, pc, pv) = _csr_args(a)
    st = np.zeros(m, np.int32)
    fi = np.zeros(m, np.int32)
    gr = np.zeros(m, np.int32)
    optr = np.zeros(m + 1, np.int64)
    cap = 256 * m + 1024
    oidx = np.zeros(cap, np.int64)
    ocoef = np.zeros(cap)
    t = L.oracle_spai_subset(n, pp, pc, pv, float(epsilon), int(alpha), int(beta), m,
                             _p(cols, C.c_int64), _p(st, C.c_int32), _p(fi, C.c_int32),
                             _p(gr, C.c_int32), _p(optr, C.c_int64), _p(oidx, C.c_int64),
                             _p(ocoef, C.c_double), cap)
    if t < 0:
        raise RuntimeError("oracle_spai_subset failed: " + L.oracle_last_error().decode())
    return st, fi, gr, optr, oidx[:t], ocoef[:t]


def score(rr, dot, den):
    st = C.c_int()
    cl = C.c_int()
    s = lib().oracle_score(rr, dot, den, C.byref(st), C.byref(cl))
    return SPAI_STATUS[st.value], s, bool(cl.value)


def from_entries(rows, cols, row, col, val):
    """reference CsrMatrix::from_entries (src/cpu.cpp:48-76): returns (ptr,
    col, val) as int64/int64/float64 arrays; raises ValueError with the
    reference's message."""
    L = lib()
    r = np.ascontiguousarray(row, dtype=np.int64)
    c = np.ascontiguousarray(col, dtype=np.int64)
    v = np.ascontiguousarray(val, dtype=np.float64)
    o = L.oracle_from_entries(rows, cols, r.size, _p(r, C.c_int64), _p(c, C.c_int64),
                              _p(v, C.c_double))
    if not o:
        raise ValueError(L.oracle_last_error().decode())
    try:
        nnz = L.oracle_csr_nnz(o)
        ptr = np.zeros(rows + 1, dtype=np.int64)
        oc = np.zeros(max(nnz, 1), dtype=np.int64)
        ov = np.zeros(max(nnz, 1), dtype=np.float64)
        L.oracle_csr_get(o, _p(ptr, C.c_int64), _p(oc, C.c_int64), _p(ov, C.c_double))
        return ptr, oc[:nnz], ov[:nnz]
    finally:
        L.oracle_csr_free(o)


def hypot_chain(v) -> float:
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().oracle_hypot_chain(_p(v, C.c_double), v.size)


@dataclass
class GmresResult:
    status: str
    converged: bool
    iterations: int
    solution: np.ndarray
    initial_original: float
    initial_relative: float
    final_original: float
    final_relative: float


def gmres(a, p, rhs, tolerance=1e-8, max_iterations=0, x0=None) -> GmresResult:
    """reference left_preconditioned_gmres (src/gmres.cpp:108-262)."""
    L = lib()
    n = a.n
    ka, (ap, ac, av) = _csr_args(a)
    kp, (pp, pc, pv) = _csr_args(p)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    sol = np.zeros(n)
    info = np.zeros(3, np.int64)
    res = np.zeros(4)
    rc = L.oracle_gmres(n, ap, ac, av, pp, pc, pv, _p(b, C.c_double),
                        None if x0a is None else _p(x0a, C.c_double), float(tolerance),
                        int(max_iterations), _p(sol, C.c_double), _p(info, C.c_int64),
                        _p(res, C.c_double))
    if rc != 0:
        raise RuntimeError(L.oracle_last_error().decode())
    return GmresResult(GMRES_STATUS[info[0]], bool(info[1]), int(info[2]), sol, *res.tolist())


@dataclass
class IrResult:
    status: str
    converged: bool
    outer_steps: int
    inner_attempts: int
    message: str
    solution: np.ndarray
    history: np.ndarray        # [h, 4]: residual_inf, residual_2, nbe, cbe
    inner: np.ndarray          # [i, 3]: status, converged, iterations
    spai_status: str | None
    p: tuple | None            # (ptr, col, val) when SPAI was constructed

    @property
    def inner_iterations(self):
        return [int(x) for x in self.inner[:, 2]]


def spai_gmres_ir(a, rhs, epsilon=0.3, alpha=None, beta=8, tolerance=1e-8, max_iterations=0,
                  max_outer_steps=20, p=None, relres_target=0.0) -> IrResult:
    """reference spai_gmres_ir / refine_with_preconditioner (src/ir.cpp:138-227)."""
    L = lib()
    n = a.n
    if alpha is None:
        alpha = -(-n // 8)
    ka, (ap, ac, av) = _csr_args(a)
    if p is not None:
        kp, (pp, pc, pv) = _csr_args(p)
    else:
        pp = pc = pv = None
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    h = L.oracle_ir(n, ap, ac, av, pp, pc, pv, _p(b, C.c_double), float(epsilon), int(alpha),
                    int(beta), float(tolerance), int(max_iterations), int(max_outer_steps),
                    float(relres_target))
    if not h:
        raise RuntimeError(L.oracle_last_error().decode())
    try:
        info = np.zeros(6, np.int64)
        L.oracle_ir_info(h, _p(info, C.c_int64))
        sol = np.zeros(n)
        if info[4] > 0:
            L.oracle_ir_solution(h, _p(sol, C.c_double))
        hist = np.zeros((int(info[4]), 4))
        if info[4] > 0:
            L.oracle_ir_history(h, _p(hist, C.c_double))
        inner = np.zeros((int(info[3]), 3), np.int64)
        if info[3] > 0:
            L.oracle_ir_inner(h, _p(inner, C.c_int64))
        pres = None
        if info[5] >= 0 and p is None:
            pn = L.oracle_ir_p_nnz(h)
            pr = np.empty(n + 1, np.int64)
            pcol = np.empty(pn, np.int64)
            pval = np.empty(pn)
            L.oracle_ir_p(h, _p(pr, C.c_int64), _p(pcol, C.c_int64), _p(pval, C.c_double))
            pres = (pr, pcol, pval)
        return IrResult(IR_STATUS[info[0]], bool(info[1]), int(info[2]), int(info[3]),
                        L.oracle_ir_message(h).decode(), sol, hist, inner,
                        SPAI_STATUS[info[5]] if info[5] >= 0 else None, pres)
    finally:
        L.oracle_ir_free(h)
