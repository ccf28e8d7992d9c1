"""Differentiable solve (2-D edge-conductivity diffusion, 256x256, n = 65,536).

L = 1/2 ||x - x*||^2 with A(theta) x = b, theta = exp(g) on edges.  The GPU
gradient (mixed FP32/FP64 forward, adjoint A^T lambda = x - x*
right-preconditioned by P^T, dL/dA = -lambda x^T on A's pattern by SDDMM) is
checked against
  * the FP64 CPU reference gradient (sparse LU for x and lambda): normwise
    relative error <= 1e-6;
  * central finite differences (h = 1e-6) on 64 random edges, to within the
    noise of the finite difference itself: |fd - g| <= 5e-3 |g| + 5e-3 max|g|
    (with exact LU solves the h = 1e-6 differences already scatter by up to
    ~1e-2 relative on small-gradient edges, because L ~ 3e6);
  * the closed form dL/dA = -lambda x^T on a tiny dense system;
and the adjoint solve itself (both precisions) against a sparse LU solve of
A^T lambda = g on the nonsymmetric 64x64 convection-diffusion system.

Both solves stop on the reference's backward-error test (eta <= 10u).  A
relative residual of 1e-12 lies below this system's FP64 floor: u ||(|A||x|)||
/ ||b|| = 2.8e-12, and LU followed by FP64 refinement stalls at 1.28e-12; the
forward test asserts that the solve reaches that floor (<= 2e-12).  The FP32
correction tolerance is the reference's FP32 default 1e-4.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c5():
    from bspai_hip import problems as P

    return P.diffusion_edges_2d(side=256)


def _loss_and_grad(handle, model, theta_np, b_np, xs_np, **opts):
    from bspai_hip import autograd as AG

    dev = torch.device("cuda:0")
    theta = torch.tensor(theta_np, device=dev, requires_grad=True)
    b = torch.tensor(b_np, device=dev)
    xs = torch.tensor(xs_np, device=dev)
    vals = AG.edge_values(model, theta)
    x = AG.solve(vals, b, handle, **opts)
    loss = 0.5 * torch.sum((x - xs) ** 2)
    loss.backward()
    return loss.item(), theta.grad.detach().cpu().numpy(), x.detach().cpu().numpy()


# Forward mixed; the adjoint in FP64: its unpreconditioned residual test has
# an FP32 floor of u32 ||A|| ||lambda|| / ||x - x*|| ~ 1.5e-3 on this system.
OPTS = dict(precision="mixed", tolerance=1e-3, max_iterations=600, relres_target=2e-12,
            adjoint=dict(precision="fp64", tolerance=1e-8, relres_target=0.0))


def test_adjoint_gradient_vs_fp64_reference(ctx, c5):
    from bspai_hip import autograd as AG

    model, theta, xs, b = c5
    h = AG.SolverHandle(model.pattern, ctx)
    loss, g, x = _loss_and_grad(h, model, theta, b, xs, **OPTS)
    gref, xref, lam = AG.reference_gradient(model, theta, b, xs)
    assert h.last_forward.status == "success" and h.last_adjoint.status == "success"
    assert np.linalg.norm(x - xref) <= 1e-9 * np.linalg.norm(xref)
    a = model.pattern
    import scipy.sparse as sp
    A = sp.csr_matrix((model.values(theta), a.col, a.ptr), shape=(a.n, a.n))
    relres = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert relres <= 2e-12, relres
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"adjoint gradient normwise rel err {rel:.3e}; forward inner "
          f"{h.last_forward.inner_iterations}, adjoint inner {h.last_adjoint.inner_iterations}")
    assert rel <= 1e-6


def test_finite_differences_64_edges(ctx, c5):
    from bspai_hip import autograd as AG

    model, theta, xs, b = c5
    h = AG.SolverHandle(model.pattern, ctx)
    fp64 = dict(precision="fp64", tolerance=1e-8, max_iterations=600)
    _, g, _ = _loss_and_grad(h, model, theta, b, xs, **fp64)
    rng = np.random.default_rng(64)
    edges = rng.choice(model.n_edges, 64, replace=False)
    dev = torch.device("cuda:0")
    bt = torch.tensor(b, device=dev)

    def L(th):
        vals = AG.edge_values(model, torch.tensor(th, device=dev))
        with torch.no_grad():
            x = AG.solve(vals, bt, h, **fp64)
        return 0.5 * float(np.sum((x.cpu().numpy() - xs) ** 2))

    step = 1e-6
    gmax = np.max(np.abs(g))
    worst = 0.0
    for e in edges:
        tp = theta.copy()
        tp[e] += step
        tm = theta.copy()
        tm[e] -= step
        fd = (L(tp) - L(tm)) / (2 * step)
        err = abs(fd - g[e])
        worst = max(worst, err / (abs(g[e]) + gmax))
        assert err <= 5e-3 * abs(g[e]) + 5e-3 * gmax, (e, fd, g[e])
    print(f"finite differences: worst |fd-g|/(|g|+max|g|) = {worst:.2e}")


def test_sddmm_closed_form(ctx):
    """dL/dvals = -lambda_i x_j on the pattern, for a dense 12x12 system."""
    from bspai_hip import autograd as AG
    from bspai_hip import problems as P

    rng = np.random.default_rng(5)
    n = 12
    M = rng.standard_normal((n, n)) + 8 * np.eye(n)
    r, c = np.nonzero(M)
    a = P._from_coo(n, r, c, M[r, c])
    h = AG.SolverHandle(a, ctx)
    dev = torch.device("cuda:0")
    vals = torch.tensor(a.val, device=dev, requires_grad=True)
    b = torch.tensor(rng.standard_normal(n), device=dev, requires_grad=True)
    xs = rng.standard_normal(n)
    x = AG.solve(vals, b, h, precision="fp64", tolerance=1e-10)
    loss = 0.5 * torch.sum((x - torch.tensor(xs, device=dev)) ** 2)
    loss.backward()
    xe = np.linalg.solve(M, b.detach().cpu().numpy())
    lam = np.linalg.solve(M.T, xe - xs)
    rows = np.repeat(np.arange(n), np.diff(a.ptr))
    expect = -lam[rows] * xe[a.col]
    assert np.allclose(vals.grad.cpu().numpy(), expect, rtol=1e-9, atol=1e-12)
    assert np.allclose(b.grad.cpu().numpy(), lam, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_adjoint_solve_nonsymmetric(ctx, precision):
    """A^T lambda = g right-preconditioned by P^T on the 64x64
    convection-diffusion matrix (A != A^T), under IR to the reference's
    backward-error test: lambda matches a sparse LU solve of A^T to 1e-10."""
    import scipy.sparse.linalg as spla

    from bspai_hip import api
    from bspai_hip import problems as P

    a = P.convection_diffusion_2d()
    m = api.Matrix(a, ctx)
    spai = api.construct_spai(m, 0.3, -(-a.n // 8), 8, precision=precision)
    assert spai.status == "success"
    g = np.random.default_rng(11).standard_normal(a.n)
    tol = 1e-8 if precision == "fp64" else 1e-4
    r = api.adjoint_solve(m.transpose(), spai.matrix(transpose=True), g, precision=precision,
                          tolerance=tol)
    assert r.status == "success" and r.converged, (r.status, r.message)
    A = a.to_scipy()
    ref = spla.spsolve(A.T.tocsc(), g)
    lam = np.asarray(r.solution)
    assert np.linalg.norm(lam - ref) <= 1e-10 * np.linalg.norm(ref)
    assert np.linalg.norm(g - A.T @ lam) <= 1e-12 * np.linalg.norm(g)
