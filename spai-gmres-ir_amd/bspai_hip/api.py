"""Python front end of the gfx950 SPAI-GMRES-IR library.

Names and argument meanings follow the reference's C++ API (include/bspai/*.hpp
in the reference tree): construct_spai, left_preconditioned_gmres,
refine_with_preconditioner, spai_gmres_ir, with the same option defaults
(SpaiOptions{epsilon=.3, alpha=5, beta=8}, GmresOptions{tolerance=1e-8,
max_iterations=0}, IrOptions{max_outer_steps=20}) and the same status names.
Every call runs on the GPU through the C ABI; host numpy arrays are copied in
and out, torch device tensors are used in place.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (DEVICE, FP64, GMRES_STATUS, HOST, IR_STATUS, MIXED, SPAI_STATUS, GmresInfo,
                   GmresOptions, IrInfo, IrOptions, SpaiInfo, SpaiOptions, check)

PRECISIONS = {"fp64": FP64, "mixed": MIXED, FP64: FP64, MIXED: MIXED}


def _addr(a):
    """Raw data address of a numpy array or a torch tensor."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def _where(a):
    """HOST or DEVICE.  A device tensor may still be being written by torch on
    its current stream, which is not the library's stream in general: wait
    for that stream before the library reads the tensor."""
    if a is None or isinstance(a, np.ndarray):
        return HOST
    if getattr(a, "is_cuda", False):
        import torch

        torch.cuda.current_stream(a.device).synchronize()
        return DEVICE
    return HOST


def _host(a, dtype):
    """The array the C ABI reads: `dtype`, contiguous.  Device tensors stay on
    the device (converted by torch on its current stream when their dtype or
    layout differs); host tensors and sequences become numpy arrays."""
    if isinstance(a, np.ndarray):
        return np.ascontiguousarray(a, dtype=dtype)
    if getattr(a, "is_cuda", False):
        import torch

        return a.detach().to(dtype=getattr(torch, np.dtype(dtype).name)).contiguous()
    if hasattr(a, "detach"):  # a host torch tensor
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a), dtype=dtype)


def _len(a) -> int:
    return len(a) if isinstance(a, np.ndarray) else int(a.numel())


def _need(a, n: int, what: str):
    """Raise ValueError unless `a` holds exactly n elements (a 1-D vector)."""
    got = _len(a)
    if got != n or (isinstance(a, np.ndarray) and a.ndim != 1) or \
            (not isinstance(a, np.ndarray) and a.dim() != 1):
        raise ValueError(f"{what}: expected a vector of {n} elements, got shape "
                         f"{tuple(a.shape)}")
    return a


class Context:
    """Device + stream (reference: the `int device` argument of the CUDA API)."""

    def __init__(self, device: int = 0, stream=None):
        L = _lib.load()
        h = C.c_void_p()
        rc = L.bspai_hip_ctx_create(device, stream, C.byref(h))
        if rc != _lib.OK:
            raise _lib.HipError(f"bspai_hip_ctx_create failed with code {rc}")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            _lib.load().bspai_hip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(_lib.load().bspai_hip_synchronize(self.h), self.h)

    def set_timing(self, enable=True):
        """Start (clearing earlier records) or stop per-kernel HIP-event timing:
        True / 1 every hot kernel, "spai" / 2 the SPAI construction only."""
        mode = 2 if enable == "spai" else int(enable)
        check(_lib.load().bspai_hip_ctx_set_timing(self.h, mode), self.h)

    def graph_stats(self):
        """(captured, replayed): Arnoldi-step graphs of this context's solves."""
        cap, rep = C.c_int64(), C.c_int64()
        check(_lib.load().bspai_hip_ctx_graph_stats(self.h, C.byref(cap), C.byref(rep)), self.h)
        return cap.value, rep.value

    def kernel_times(self):
        """Per-kernel totals since timing was enabled:
        {label: dict(launches, total_ms, avg_us, bytes, flops)}."""
        L = _lib.load()
        count = C.c_int(0)
        check(L.bspai_hip_ctx_kernel_times(self.h, None, 0, C.byref(count)), self.h)
        buf = (_lib.KernelTime * max(1, count.value))()
        got = C.c_int(0)
        check(L.bspai_hip_ctx_kernel_times(self.h, buf, len(buf), C.byref(got)), self.h)
        out = {}
        for i in range(min(got.value, len(buf))):
            k = buf[i]
            out[k.name.decode()] = dict(launches=int(k.launches), total_ms=k.total_ms,
                                        avg_us=1e3 * k.total_ms / max(1, k.launches),
                                        bytes=k.bytes, flops=k.flops)
        return out


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def csr_from_entries(rows: int, cols: int, row, col, val, ctx: Context | None = None):
    """Canonical CSR from coordinate entries on the GPU (reference
    CsrMatrix::from_entries, src/cpu.cpp:48-76): sorted by (row, column),
    duplicates summed in input order, exact zeros dropped.  Returns (ptr, col,
    val): int32/int32/float64 numpy arrays for host inputs, torch tensors on
    the inputs' device for device inputs.  Raises ValueError with the
    reference's message for an index outside the matrix, a non-finite value
    or a non-finite duplicate sum."""
    ctx = ctx or default_context()
    r = _host(row, np.int64)
    c = _host(col, np.int64)
    v = _host(val, np.float64)
    count = _len(v)
    _need(r, count, "row")
    _need(c, count, "col")
    _need(v, count, "val")
    where = _where(v)
    if where == DEVICE and not (getattr(r, "is_cuda", False) and getattr(c, "is_cuda", False)):
        raise ValueError("row, col and val must all be host arrays or all device tensors")
    if where == DEVICE:
        import torch

        dev = v.device
        ptr = torch.empty(rows + 1, dtype=torch.int32, device=dev)
        oc = torch.empty(max(count, 1), dtype=torch.int32, device=dev)
        ov = torch.empty(max(count, 1), dtype=torch.float64, device=dev)
    else:
        ptr = np.empty(rows + 1, np.int32)
        oc = np.empty(max(count, 1), np.int32)
        ov = np.empty(max(count, 1))
    nnz = C.c_int64()
    check(_lib.load().bspai_hip_csr_from_entries(
        ctx.h, rows, cols, count, _addr(r), _addr(c), _addr(v), where, _addr(ptr), _addr(oc),
        _addr(ov), where, C.byref(nnz)), ctx.h)
    k = nnz.value
    return ptr, oc[:k], ov[:k]


class MatrixMarketError(RuntimeError):
    """A malformed Matrix Market text (the reference reader's std::runtime_error)."""


def parse_matrix_market(text, threads: int = 0, gpu: bool = False, ctx: Context | None = None):
    """Coordinate entries of a Matrix Market text, parsed with the rules and
    messages of the reference's read_matrix_market (src/cpu.cpp:105-161): on
    the host by the library's multi-threaded reader (bspai_hip_mm_parse), or
    with gpu=True on the GPU (bspai_hip_mm_parse_gpu, same rules and bits).

    Returns (rows, cols, row, col, val): zero-based int64 / float64 numpy
    arrays in file order (a symmetric file's off-diagonal entries followed by
    their mirror), not yet canonicalized.  Raises MatrixMarketError with the
    reference's message for a malformed file."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    L = _lib.load()
    if gpu:
        ctx = ctx or default_context()
        h = C.c_void_p()
        msg = C.create_string_buffer(256)
        rc = L.bspai_hip_mm_parse_gpu(ctx.h, data, len(data), C.byref(h), msg, len(msg))
        if rc == _lib.EFORMAT:
            raise MatrixMarketError(msg.value.decode())
        check(rc, ctx.h)
        try:
            rows, cols, count = C.c_int64(), C.c_int64(), C.c_int64()
            L.bspai_hip_mm_gpu_info(h, C.byref(rows), C.byref(cols), C.byref(count), None, None)
            r = np.empty(count.value, np.int64)
            c = np.empty(count.value, np.int64)
            v = np.empty(count.value, np.float64)
            check(L.bspai_hip_mm_gpu_copy(ctx.h, h, r.ctypes.data, c.ctypes.data, v.ctypes.data,
                                          HOST), ctx.h)
        finally:
            L.bspai_hip_mm_gpu_destroy(h)
        return rows.value, cols.value, r, c, v
    h = C.c_void_p()
    msg = C.create_string_buffer(256)
    rc = L.bspai_hip_mm_parse(data, len(data), int(threads), C.byref(h), msg, len(msg))
    if rc == _lib.EFORMAT:
        raise MatrixMarketError(msg.value.decode())
    if rc != _lib.OK:
        raise (MemoryError if rc == _lib.ENOMEM else ValueError)(msg.value.decode())
    try:
        rows, cols, count = C.c_int64(), C.c_int64(), C.c_int64()
        L.bspai_hip_mm_info(h, C.byref(rows), C.byref(cols), C.byref(count), None, None)
        r = np.empty(count.value, np.int64)
        c = np.empty(count.value, np.int64)
        v = np.empty(count.value, np.float64)
        check(L.bspai_hip_mm_entries(h, r.ctypes.data, c.ctypes.data, v.ctypes.data), None)
    finally:
        L.bspai_hip_mm_destroy(h)
    return rows.value, cols.value, r, c, v


def read_matrix_market(source, ctx: Context | None = None, gpu: bool = True):
    """Canonical CSR of a Matrix Market file (a path) or text (str/bytes with
    a newline): parsed on the GPU (gpu=False: on the host,
    parse_matrix_market), sorted and summed on the GPU (csr_from_entries) --
    reference read_matrix_market (include/bspai/cpu.hpp:47).  Returns (rows,
    cols, ptr, col, val); raises MatrixMarketError or ValueError with the
    reference's messages."""
    if isinstance(source, (bytes, bytearray)) or (isinstance(source, str) and "\n" in source):
        text = source
    else:
        with open(source, "rb") as f:
            text = f.read()
    rows, cols, r, c, v = parse_matrix_market(text, gpu=gpu, ctx=ctx)
    ptr, col, val = csr_from_entries(rows, cols, r, c, v, ctx)
    return rows, cols, ptr, col, val


class Matrix:
    """Device CSR (int32 structure, FP64 values); optional block-diagonal batch."""

    @classmethod
    def from_entries(cls, n: int, row, col, val, ctx: Context | None = None) -> "Matrix":
        """n x n matrix from coordinate entries, canonicalized on the GPU."""
        ptr, c, v = csr_from_entries(n, n, row, col, val, ctx)
        return cls((ptr, c, v), ctx)

    def __init__(self, a, ctx: Context | None = None, systems=None, _handle=None):
        self.ctx = ctx or default_context()
        self._keep = []
        if _handle is not None:
            self.h = _handle
        else:
            ptr, col, val = (a.ptr, a.col, a.val) if hasattr(a, "ptr") else a
            ptr = _host(ptr, np.int32)
            col = _host(col, np.int32)
            val = _host(val, np.float64)
            n = _len(ptr) - 1
            nnz = _len(val)
            if n < 1:
                raise ValueError("ptr must hold n + 1 >= 2 row offsets")
            _need(col, nnz, "col")
            where = _where(val)
            if (_where(ptr) != where) or (_where(col) != where):
                raise ValueError("ptr, col and val must all be host arrays or all device tensors")
            nsys, offs = 1, None
            if systems is not None:
                offs = np.ascontiguousarray(systems, dtype=np.int64)
                nsys = len(offs) - 1
            h = C.c_void_p()
            check(_lib.load().bspai_hip_matrix_create(
                self.ctx.h, n, nnz, _addr(ptr), _addr(col), _addr(val), where, nsys,
                None if offs is None else offs.ctypes.data, C.byref(h)), self.ctx.h)
            self.h = h
        n = C.c_int64()
        nnz = C.c_int64()
        ns = C.c_int()
        _lib.load().bspai_hip_matrix_info(self.h, C.byref(n), C.byref(nnz), C.byref(ns))
        self.n, self.nnz, self.nsys = n.value, nnz.value, ns.value

    def close(self):
        if getattr(self, "h", None):
            _lib.load().bspai_hip_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def transpose(self) -> "Matrix":
        h = C.c_void_p()
        check(_lib.load().bspai_hip_matrix_transpose(self.ctx.h, self.h, C.byref(h)), self.ctx.h)
        return Matrix(None, self.ctx, _handle=h)

    def set_values(self, val):
        v = _need(_host(val, np.float64), self.nnz, "values")
        check(_lib.load().bspai_hip_matrix_set_values(self.ctx.h, self.h, _addr(v), _where(v)),
              self.ctx.h)

    def download(self):
        ptr = np.empty(self.n + 1, np.int32)
        col = np.empty(max(self.nnz, 1), np.int32)
        val = np.empty(max(self.nnz, 1))
        check(_lib.load().bspai_hip_matrix_download(self.ctx.h, self.h, ptr.ctypes.data,
                                                    col.ctypes.data, val.ctypes.data),
              self.ctx.h)
        return ptr, col[: self.nnz], val[: self.nnz]

    def multiply(self, x, lanes: int = 0):
        """y = A x in FP64 (reference CsrMatrix::multiply, src/cpu.cpp:80-94)."""
        xs = _need(_host(x, np.float64), self.n, "x")
        if _where(xs) == DEVICE:
            import torch

            y = torch.empty(self.n, dtype=torch.float64, device=xs.device)
        else:
            y = np.empty(self.n)
        check(_lib.load().bspai_hip_spmv(self.ctx.h, self.h, _addr(xs), _addr(y), _where(xs),
                                         lanes), self.ctx.h)
        self.ctx.synchronize()
        return y

    def multiply_f32(self, x, lanes: int = 0):
        xs = _need(_host(x, np.float32), self.n, "x")
        if _where(xs) == DEVICE:
            import torch

            y = torch.empty(self.n, dtype=torch.float32, device=xs.device)
        else:
            y = np.empty(self.n, np.float32)
        check(_lib.load().bspai_hip_spmv_f32(self.ctx.h, self.h, _addr(xs), _addr(y),
                                             _where(xs), lanes), self.ctx.h)
        self.ctx.synchronize()
        return y

    def sddmm(self, u, v, alpha: float = 1.0):
        """out[q] = alpha * u[row(q)] * v[col(q)] on the stored pattern."""
        us = _need(_host(u, np.float64), self.n, "u")
        vs = _need(_host(v, np.float64), self.n, "v")
        if _where(us) != _where(vs):
            raise ValueError("u and v must both be host arrays or both device tensors")
        if _where(us) == DEVICE:
            import torch

            out = torch.empty(self.nnz, dtype=torch.float64, device=us.device)
        else:
            out = np.empty(max(self.nnz, 1))
        check(_lib.load().bspai_hip_sddmm(self.ctx.h, self.h, float(alpha), _addr(us), _addr(vs),
                                          _addr(out), _where(us)), self.ctx.h)
        self.ctx.synchronize()
        return out if _where(us) == DEVICE else out[: self.nnz]


@dataclass
class SpaiResult:
    status: str
    message: str
    failing_column: int
    p_nnz: int
    pattern_nnz: int
    fits: int
    growths: int
    clamps: int
    passes: int
    build_ms: float
    handle: object = field(default=None, repr=False)
    ctx: object = field(default=None, repr=False)
    n: int = 0

    def preconditioner(self):
        """Canonical P = N^T D as host (ptr int64, col int64, val float64)."""
        ptr = np.empty(self.n + 1, np.int32)
        col = np.empty(max(self.p_nnz, 1), np.int32)
        val = np.empty(max(self.p_nnz, 1))
        check(_lib.load().bspai_hip_spai_preconditioner(self.ctx.h, self.handle, ptr.ctypes.data,
                                                        col.ctypes.data, val.ctypes.data),
              self.ctx.h)
        return ptr.astype(np.int64), col[: self.p_nnz].astype(np.int64), val[: self.p_nnz]

    def columns(self):
        n = self.n
        st = np.empty(n, np.int32)
        fi = np.empty(n, np.int32)
        gr = np.empty(n, np.int32)
        cl = np.empty(n, np.int32)
        rs = np.empty(n)
        pv = np.empty(n)
        check(_lib.load().bspai_hip_spai_columns(self.ctx.h, self.handle, st.ctypes.data,
                                                 fi.ctypes.data, gr.ctypes.data, cl.ctypes.data,
                                                 rs.ctypes.data, pv.ctypes.data), self.ctx.h)
        return dict(status=st, fits=fi, growths=gr, clamps=cl, residual=rs, pivot_ratio=pv)

    def patterns(self):
        ptr = np.empty(self.n + 1, np.int64)
        idx = np.empty(max(self.pattern_nnz, 1), np.int32)
        coef = np.empty(max(self.pattern_nnz, 1))
        check(_lib.load().bspai_hip_spai_patterns(self.ctx.h, self.handle, ptr.ctypes.data,
                                                  idx.ctypes.data, coef.ctypes.data), self.ctx.h)
        return ptr, idx[: self.pattern_nnz].astype(np.int64), coef[: self.pattern_nnz]

    def scaling(self):
        d = np.empty(self.n)
        check(_lib.load().bspai_hip_spai_scaling(self.ctx.h, self.handle, d.ctypes.data),
              self.ctx.h)
        return d

    def matrix(self, transpose: bool = False) -> Matrix:
        h = C.c_void_p()
        check(_lib.load().bspai_hip_spai_matrix(self.ctx.h, self.handle, int(transpose),
                                                C.byref(h)), self.ctx.h)
        return Matrix(None, self.ctx, _handle=h)

    def __del__(self):
        try:
            if self.handle:
                _lib.load().bspai_hip_spai_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def construct_spai(a: Matrix, epsilon: float = 0.3, alpha: int = 5, beta: int = 8,
                   precision="fp64") -> SpaiResult:
    """Adaptive SPAI on the GPU (reference construct_spai, src/spai.cpp:314-374)."""
    ctx = a.ctx
    opt = SpaiOptions(float(epsilon), int(alpha), int(beta), PRECISIONS[precision])
    h = C.c_void_p()
    check(_lib.load().bspai_hip_spai_build(ctx.h, a.h, C.byref(opt), C.byref(h)), ctx.h)
    info = SpaiInfo()
    _lib.load().bspai_hip_spai_get_info(h, C.byref(info))
    return SpaiResult(SPAI_STATUS[info.status], info.message.decode(), info.failing_column,
                      info.p_nnz, info.pattern_nnz, info.fits, info.growths, info.clamps,
                      info.passes, info.build_ms, h, ctx, info.n)


def evaluate_spai_score(residual_squared: float, dot: float, denominator_squared: float,
                        ctx: Context | None = None):
    """Restricted score on the device (reference evaluate_spai_score, src/spai.cpp:273-306)."""
    ctx = ctx or default_context()
    st = C.c_int()
    cl = C.c_int()
    sc = C.c_double()
    check(_lib.load().bspai_hip_score(ctx.h, residual_squared, dot, denominator_squared,
                                      C.byref(st), C.byref(cl), C.byref(sc)), ctx.h)
    return SPAI_STATUS[st.value], sc.value, bool(cl.value)


@dataclass
class GmresResult:
    status: str
    converged: bool
    iterations: int
    solution: object
    initial_original_residual: float
    initial_preconditioned_relative: float
    final_original_residual: float
    final_preconditioned_relative: float


def left_preconditioned_gmres(a: Matrix, p: Matrix, rhs, tolerance: float = 1e-8,
                              max_iterations: int = 0, initial_guess=None):
    """FP64 GMRES on P A x = P b per system (reference src/gmres.cpp:108-262).

    Returns one GmresResult (single system) or a list (batched matrix)."""
    ctx = a.ctx
    b = _need(_host(rhs, np.float64), a.n, "rhs")
    x0 = None if initial_guess is None else _need(_host(initial_guess, np.float64), a.n,
                                                  "initial_guess")
    where = _where(b)
    if x0 is not None and _where(x0) != where:
        raise ValueError("rhs and initial_guess must both be host arrays or both device tensors")
    if where == DEVICE:
        import torch

        x = torch.empty(a.n, dtype=torch.float64, device=b.device)
    else:
        x = np.empty(a.n)
    infos = (GmresInfo * a.nsys)()
    opt = GmresOptions(float(tolerance), int(max_iterations))
    check(_lib.load().bspai_hip_gmres(ctx.h, a.h, p.h, _addr(b), _addr(x0), C.byref(opt),
                                      _addr(x), where, infos), ctx.h)
    out = [GmresResult(GMRES_STATUS[i.status], bool(i.converged), int(i.iterations), x,
                       i.initial_original, i.initial_relative, i.final_original, i.final_relative)
           for i in infos]
    return out[0] if a.nsys == 1 else out


@dataclass
class IrResult:
    status: str
    converged: bool
    construction_performed: bool
    spai_status: str
    outer_steps: int
    inner_attempts: int
    inner_statuses: list
    inner_iterations: list
    history: np.ndarray  # [steps+1, 4]: residual_inf, residual_2, nbe, cbe
    residual_inf: float
    residual_2: float
    normwise_backward_error: float
    componentwise_backward_error: float
    spai_ms: float
    solve_ms: float
    message: str
    solution: object = None


def _ir_info(i: IrInfo, x) -> IrResult:
    k = int(i.inner_attempts)
    hist = np.array([[i.history[h][q] for q in range(4)] for h in range(i.history_len)])
    return IrResult(IR_STATUS[i.status], bool(i.converged), bool(i.construction_performed),
                    SPAI_STATUS[i.spai_status], int(i.outer_steps), k,
                    [GMRES_STATUS[i.inner_status[j]] for j in range(k)],
                    [int(i.inner_iterations[j]) for j in range(k)], hist, i.residual_inf,
                    i.residual_2, i.normwise_backward_error, i.componentwise_backward_error,
                    i.spai_ms, i.solve_ms, i.message.decode(), x)


def ir_options(epsilon=0.3, alpha=0, beta=8, tolerance=1e-8, max_iterations=0,
               max_outer_steps=20, precision="fp64", relres_target=0.0) -> IrOptions:
    return IrOptions(float(epsilon), int(alpha), int(beta), float(tolerance),
                     int(max_iterations), int(max_outer_steps), PRECISIONS[precision],
                     float(relres_target))


def _x_like(b, n):
    if _where(b) == DEVICE:
        import torch

        return torch.empty(n, dtype=torch.float64, device=b.device)
    return np.empty(n)


def spai_gmres_ir(a: Matrix, rhs, **kw):
    """SPAI construction + IR (reference spai_gmres_ir, src/ir.cpp:206-227).

    Keywords: epsilon, alpha (0 = ceil(n/8), as the CLI), beta, tolerance,
    max_iterations, max_outer_steps, precision ("fp64" | "mixed"),
    relres_target.  Returns an IrResult (or a list for a batched matrix)."""
    ctx = a.ctx
    b = _need(_host(rhs, np.float64), a.n, "rhs")
    x = _x_like(b, a.n)
    infos = (IrInfo * a.nsys)()
    opt = ir_options(**kw)
    check(_lib.load().bspai_hip_spai_gmres_ir(ctx.h, a.h, _addr(b), C.byref(opt), _addr(x),
                                              _where(b), infos), ctx.h)
    out = [_ir_info(i, x) for i in infos]
    return out[0] if a.nsys == 1 else out


def refine_with_preconditioner(a: Matrix, p, rhs, **kw):
    """IR with a given preconditioner (reference refine_with_preconditioner,
    src/ir.cpp:138-204).  p: a Matrix, or a SpaiResult (its P is used)."""
    ctx = a.ctx
    pm = p.matrix() if isinstance(p, SpaiResult) else p
    b = _need(_host(rhs, np.float64), a.n, "rhs")
    x = _x_like(b, a.n)
    infos = (IrInfo * a.nsys)()
    opt = ir_options(**kw)
    check(_lib.load().bspai_hip_ir_solve(ctx.h, a.h, pm.h, _addr(b), C.byref(opt), _addr(x),
                                         _where(b), infos), ctx.h)
    out = [_ir_info(i, x) for i in infos]
    return out[0] if a.nsys == 1 else out


def adjoint_solve(at: Matrix, pt: Matrix, g, **kw):
    """Solve A^T lambda = g, right-preconditioned by P^T, under IR."""
    ctx = at.ctx
    gs = _need(_host(g, np.float64), at.n, "g")
    lam = _x_like(gs, at.n)
    infos = (IrInfo * at.nsys)()
    opt = ir_options(**kw)
    check(_lib.load().bspai_hip_adjoint_solve(ctx.h, at.h, pt.h, _addr(gs), C.byref(opt),
                                              _addr(lam), _where(gs), infos), ctx.h)
    out = [_ir_info(i, lam) for i in infos]
    return out[0] if at.nsys == 1 else out
