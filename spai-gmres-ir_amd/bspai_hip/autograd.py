"""Differentiable sparse solve x = A(values)^-1 b on the GPU.

forward:  SPAI construction of P ~ A^-1 and SPAI-GMRES-IR for A x = b
          (mixed FP32/FP64 or plain FP64), all through the C ABI.
backward: the adjoint system A^T lambda = dL/dx, solved by the same IR with
          GMRES right-preconditioned by P^T (A P ~ I  =>  A^T P^T ~ I), then
          dL/dA = -lambda x^T restricted to A's stored pattern (SDDMM kernel)
          and dL/db = lambda.

Gradients flow to the CSR `values` tensor; any differentiable map theta ->
values (e.g. problems.EdgeDiffusion) composes with it through torch autograd.
"""
from __future__ import annotations

import numpy as np
import torch

from . import api


_SPAI_KEYS = ("epsilon", "alpha", "beta", "precision")


def _spai_options(opts: dict, n: int) -> dict:
    """SPAI parameters of a solve (alpha 0 = ceil(n/8), as the reference CLI)."""
    return dict(epsilon=opts.get("epsilon", 0.3), alpha=opts.get("alpha", 0) or -(-n // 8),
                beta=opts.get("beta", 8), precision=opts.get("precision", "fp64"))


class SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, handle, opts):
        m = handle.matrix
        m.set_values(values.detach().contiguous())
        fwd = {k: v for k, v in opts.items() if k != "adjoint"}
        # construct_spai + refine_with_preconditioner is the reference's
        # spai_gmres_ir (src/ir.cpp:206-227); P is kept for the backward pass.
        sk = _spai_options(fwd, m.n)
        spai = api.construct_spai(m, **sk)
        if spai.status != "success":
            raise RuntimeError(f"forward SPAI failed: {spai.status}: {spai.message}")
        ir = {k: v for k, v in fwd.items() if k not in ("epsilon", "alpha", "beta")}
        res = api.refine_with_preconditioner(m, spai, b.detach().contiguous(), **ir)
        if res.status != "success":
            raise RuntimeError(f"forward solve failed: {res.status}: {res.message}")
        handle.last_forward = res
        x = res.solution
        ctx.handle = handle
        ctx.opts = opts
        # the solve runs in FP64 (the library converts other dtypes); the
        # gradients are returned in the inputs' own dtypes
        ctx.dtypes = (values.dtype, b.dtype)
        ctx.spai = (spai, sk)
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        h = ctx.handle
        m = h.matrix
        opts = {k: v for k, v in ctx.opts.items() if k != "adjoint"}
        opts.update(ctx.opts.get("adjoint") or {})
        # P of the values used in forward: the forward's construction when the
        # adjoint asks for the same SPAI parameters, else a new one.
        spai, sk = ctx.spai
        if _spai_options(opts, m.n) != sk:
            spai = api.construct_spai(m, **_spai_options(opts, m.n))
        if spai.status != "success":
            raise RuntimeError(f"backward SPAI failed: {spai.status}")
        at = m.transpose()
        pt = spai.matrix(transpose=True)
        kw = {k: v for k, v in opts.items() if k not in ("epsilon", "alpha", "beta")}
        adj = api.adjoint_solve(at, pt, grad_x.contiguous(), **kw)
        if adj.status != "success":
            raise RuntimeError(f"adjoint solve failed: {adj.status}: {adj.message}")
        lam = adj.solution
        h.last_adjoint = adj
        grad_vals = m.sddmm(lam, x, alpha=-1.0)
        vdt, bdt = ctx.dtypes
        return grad_vals.to(vdt), lam.to(bdt), None, None


class SolverHandle:
    """Holds the device matrix (fixed pattern, values replaced per call)."""

    def __init__(self, pattern, ctx: api.Context | None = None):
        self.matrix = api.Matrix(pattern, ctx)
        self.last_forward = None
        self.last_adjoint = None


def solve(values: torch.Tensor, b: torch.Tensor, handle: SolverHandle, **opts) -> torch.Tensor:
    """x = A(values)^-1 b, differentiable in values and b.

    opts: spai_gmres_ir keywords (epsilon, beta, tolerance, max_iterations,
    max_outer_steps, precision, relres_target), used by both solves, plus
    `adjoint`: a dict of keywords that override them for the adjoint solve.
    The adjoint's stopping test is on the unpreconditioned residual
    ||g - A^T lambda|| / ||g||, whose FP32 floor is u32 ||A|| ||lambda|| / ||g||:
    ill-conditioned systems may need the adjoint in FP64, or a looser FP32
    tolerance, where the forward solve runs mixed."""
    return SparseSolve.apply(values, b, handle, opts)


def edge_values(model, theta: torch.Tensor) -> torch.Tensor:
    """CSR values of A(theta) for a problems.EdgeDiffusion model (differentiable)."""
    dev = theta.device
    s_val = torch.as_tensor(model.s_val, device=dev)
    s_edge = torch.as_tensor(model.s_edge, device=dev)
    s_coef = torch.as_tensor(model.s_coef, device=dev, dtype=theta.dtype)
    out = torch.zeros(model.pattern.nnz, dtype=theta.dtype, device=dev)
    return out.index_add(0, s_val, s_coef * theta[s_edge])


def reference_gradient(model, theta: np.ndarray, b: np.ndarray, x_star: np.ndarray):
    """FP64 CPU gradient of L = 1/2 ||x - x*||^2 w.r.t. theta (direct sparse LU):
    A x = b, A^T lam = x - x*, dL/dvals = -lam_i x_j, dL/dtheta = S^T dL/dvals."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    vals = model.values(theta)
    p = model.pattern
    A = sp.csr_matrix((vals, p.col, p.ptr), shape=(p.n, p.n))
    lu = spla.splu(A.tocsc())
    x = lu.solve(b)
    lam = lu.solve(x - x_star, trans="T")
    rows = np.repeat(np.arange(p.n), np.diff(p.ptr))
    gvals = -lam[rows] * x[p.col]
    return model.grad_theta(gvals), x, lam
