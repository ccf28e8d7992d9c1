"""GPU parity of the HIP path against the reference's golden outputs and the
CPU oracle (bit-exact where the arithmetic is sequential, tolerance-bounded
where the GPU reduces in parallel).

Bars (stated per test):
  * SpMV with lanes=1, SPAI FP64 (scaling, patterns, coefficients, P), IR
    diagnostics of an identical iterate: bit-exact.
  * GMRES / IR: status equal, inner iteration counts within +-1 of the
    reference, final backward error <= 10*2^-53, solution within 1e-10
    relative of the reference's.
  * Mixed precision: P entries within 1e-5 relative of the FP64 P on the common
    pattern (pattern differences counted), final rel-res <= 1e-12.
"""
import math

import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def _problems():
    from bspai_hip import problems as P

    return P


def test_context_and_abi(ctx):
    import bspai_hip

    assert bspai_hip.load().bspai_hip_abi_version() == 1
    ctx.synchronize()


@pytest.mark.parametrize("lanes", [1, 2, 8, 64])
def test_spmv_c1(ctx, lanes):
    """Row-sequential SpMV is bit-identical to CsrMatrix::multiply (src/cpu.cpp:80-94);
    grouped lanes stay within the reference's row-length bound
    (tests/test_cuda_spmv.cpp:26-39): |y - y_ref| <= (row nnz) * u * sum |a||x|."""
    import bspai_hip

    P = _problems()
    a = P.convection_diffusion_2d()
    rng = np.random.default_rng(7)
    x = rng.standard_normal(a.n)
    m = bspai_hip.Matrix(a, ctx)
    y = m.multiply(x, lanes=lanes)
    # sequential CPU product in the reference's order
    ref = np.zeros(a.n)
    for r in range(a.n):
        s = 0.0
        for q in range(a.ptr[r], a.ptr[r + 1]):
            s += a.val[q] * x[a.col[q]]
        ref[r] = s
    if lanes == 1:
        assert np.array_equal(bits(y), bits(ref))
    else:
        absprod = np.abs(a.to_scipy()) @ np.abs(x)
        rowlen = np.diff(a.ptr)
        assert np.all(np.abs(y - ref) <= rowlen * 2.0 ** -53 * absprod * 2)


def test_spmv_exact_dyadic(ctx):
    """Dyadic values give exact sums in any order (tests/test_cuda_spmv.cpp:73-89)."""
    import bspai_hip

    P = _problems()
    a = P.band(1000, 8.0)
    x = np.ldexp(1.0, -(np.arange(1000) % 7))
    y = bspai_hip.Matrix(a, ctx).multiply(x, lanes=8)
    assert np.array_equal(y, a.multiply(x))


def test_spai_c1_bit_exact(ctx, c1_golden):
    """GPU FP64 SPAI on the 64x64 stencil equals the reference's construct_spai bit for bit."""
    import bspai_hip

    P = _problems()
    a = P.convection_diffusion_2d()
    m = bspai_hip.Matrix(a, ctx)
    s = bspai_hip.construct_spai(m, epsilon=0.3, alpha=-(-a.n // 8), beta=8)
    assert s.status == "success"
    cols = s.columns()
    assert np.array_equal(cols["fits"], c1_golden["fits"])
    assert np.array_equal(cols["growths"], c1_golden["growths"])
    assert np.array_equal(cols["clamps"], c1_golden["clamps"])
    assert np.array_equal(bits(cols["residual"]), bits(c1_golden["residual"]))
    assert np.array_equal(bits(s.scaling()), bits(c1_golden["scaling"]))
    ptr, col, val = s.preconditioner()
    assert np.array_equal(ptr, c1_golden["p_ptr"])
    assert np.array_equal(col, c1_golden["p_col"])
    assert np.array_equal(bits(val), bits(c1_golden["p_val"]))


@pytest.fixture(scope="module")
def power_law_small():
    """Power-law matrix of the uneven-columns workload at n = 1000 (eps 0.2),
    with the oracle's full SPAI: its columns reach q ~ 100 and shadows of
    hundreds of rows, so most of them overflow the lane kernel's bucket and
    are fitted by the workgroup-per-column kernel."""
    from oracle import pyoracle as O

    P = _problems()
    a = P.power_law(n=1000)
    return a, O.construct_spai(a, 0.2, -(-a.n // 8), 8)


def test_spai_block_kernel_bit_exact(ctx, power_law_small):
    import bspai_hip

    a, ref = power_law_small
    m = bspai_hip.Matrix(a, ctx)
    s = bspai_hip.construct_spai(m, epsilon=0.2, alpha=-(-a.n // 8), beta=8)
    assert s.status == ref.status == "success"
    assert s.passes >= 2  # the block kernel ran
    cols = s.columns()
    assert np.array_equal(cols["fits"], ref.fits)
    assert np.array_equal(cols["growths"], ref.growths)
    assert np.array_equal(cols["clamps"], ref.clamps)
    assert np.array_equal(bits(cols["residual"]), bits(ref.residual))
    ptr, col, val = s.preconditioner()
    assert np.array_equal(ptr, ref.p_ptr)
    assert np.array_equal(col, ref.p_col)
    assert np.array_equal(bits(val), bits(ref.p_val))


def test_spai_block_kernel_fp32_and_mixed_ir(ctx, power_law_small):
    """FP32 build of the same matrix.  Row k of P holds column k's fit, so
    rows are compared: where the adaptive pattern is the oracle's, the
    coefficients agree to 1e-5 (normwise per row); rows whose FP32 selection
    took another path are counted and reported (many growth steps with close
    scores make such divergence common on this matrix).  Then mixed IR
    reaches rel-res 1e-10, the uneven-columns workload's target."""
    import bspai_hip

    a, ref = power_law_small
    m = bspai_hip.Matrix(a, ctx)
    s = bspai_hip.construct_spai(m, epsilon=0.2, alpha=-(-a.n // 8), beta=8, precision="mixed")
    assert s.status == "success" and s.passes >= 2
    ptr, col, val = s.preconditioner()
    same, worst = 0, 0.0
    for k in range(a.n):
        g = slice(ptr[k], ptr[k + 1])
        r = slice(ref.p_ptr[k], ref.p_ptr[k + 1])
        if np.array_equal(col[g], ref.p_col[r]):
            same += 1
            worst = max(worst, np.linalg.norm(val[g] - ref.p_val[r]) / np.linalg.norm(ref.p_val[r]))
    print(f"FP32 SPAI: {a.n - same} of {a.n} columns took another adaptive pattern; "
          f"worst normwise error on the others {worst:.2e}")
    assert same > 0 and worst <= 1e-5
    b = _problems().equal_rhs(a.n)
    r = bspai_hip.refine_with_preconditioner(m, s, b, precision="mixed", tolerance=1e-4,
                                             relres_target=1e-10)
    x = np.asarray(r.solution)
    A = a.to_scipy()
    assert r.status == "success"
    assert np.linalg.norm(b - A @ x) <= 1e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("kernel", ["wave", "lane"])
@pytest.mark.parametrize("side", [0, 32])
def test_spai_fp32_kernels_vs_fp64(ctx, monkeypatch, kernel, side):
    """Both FP32 first-bucket kernels (wavefront per column, lane per column),
    forced in turn, on the 64x64 stencil and the 32^3 convection-diffusion
    matrix: every column's adaptive pattern equals the FP64 build's (itself
    bit-exact with the oracle) and the coefficients agree to 1e-5 (max-norm
    per row of P, relative)."""
    import bspai_hip

    P = _problems()
    a = P.convection_diffusion_2d() if side == 0 else P.convection_diffusion_3d(side)
    m = bspai_hip.Matrix(a, ctx)
    alpha = -(-a.n // 8)
    ref = bspai_hip.construct_spai(m, epsilon=0.3, alpha=alpha, beta=8)
    monkeypatch.setenv("BSPAI_SPAI_FP32_KERNEL", kernel)
    s = bspai_hip.construct_spai(m, epsilon=0.3, alpha=alpha, beta=8, precision="mixed")
    assert ref.status == "success" and s.status == "success"
    assert s.fits == ref.fits and s.growths == ref.growths
    gp, gc, gv = ref.preconditioner()
    ptr, col, val = s.preconditioner()
    worst, diff_rows = 0.0, 0
    for r in range(a.n):
        c32, v32 = col[ptr[r]:ptr[r + 1]], val[ptr[r]:ptr[r + 1]]
        c64, v64 = gc[gp[r]:gp[r + 1]], gv[gp[r]:gp[r + 1]]
        if not np.array_equal(c32, c64):
            diff_rows += 1
            continue
        worst = max(worst, np.max(np.abs(v32 - v64)) / np.max(np.abs(v64)))
    print(f"{kernel} kernel, n={a.n}: {diff_rows} differing patterns, worst {worst:.2e}")
    assert diff_rows == 0 and worst <= 1e-5


@pytest.mark.parametrize("n", [16, 64, 256, 1024])
@pytest.mark.parametrize("diag", [8, 5])
def test_spai_band_bit_exact(ctx, band_golden, n, diag):
    """Banded family of the reference's sparse GPU SPAI test (tests/test_cuda_spai_sparse.cpp:7)."""
    import bspai_hip

    P = _problems()
    a = P.band(n, float(diag))
    g = {k[len(f"n{n}_d{diag}_"):]: v for k, v in band_golden.items()
         if k.startswith(f"n{n}_d{diag}_")}
    s = bspai_hip.construct_spai(bspai_hip.Matrix(a, ctx), epsilon=0.3, alpha=5, beta=8)
    assert s.status == str(g["spai_status"])
    cols = s.columns()
    assert np.array_equal(cols["fits"], g["fits"])
    assert np.array_equal(cols["growths"], g["growths"])
    ptr, col, val = s.preconditioner()
    assert np.array_equal(ptr, g["p_ptr"])
    assert np.array_equal(col, g["p_col"])
    assert np.array_equal(bits(val), bits(g["p_val"]))


def test_spai_tiny_closed_form(ctx):
    """[[4,1,0],[2,5,1],[0,1,3]] at eps 1e-12, alpha 2: P is the exact inverse
    [[14,-3,1],[-6,12,-4],[2,-4,18]]/50 (reference tests/test_spai.cpp:139-168)."""
    import bspai_hip

    P = _problems()
    a = P._from_coo(3, np.array([0, 0, 1, 1, 1, 2, 2]), np.array([0, 1, 0, 1, 2, 1, 2]),
                    np.array([4.0, 1, 2, 5, 1, 1, 3]))
    s = bspai_hip.construct_spai(bspai_hip.Matrix(a, ctx), epsilon=1e-12, alpha=2, beta=8)
    assert s.status == "success"
    ptr, col, val = s.preconditioner()
    dense = np.zeros((3, 3))
    for r in range(3):
        dense[r, col[ptr[r]:ptr[r + 1]]] = val[ptr[r]:ptr[r + 1]]
    exact = np.array([[14, -3, 1], [-6, 12, -4], [2, -4, 18]]) / 50.0
    assert np.allclose(dense, exact, rtol=0, atol=768 * 2.0 ** -52)


@pytest.mark.parametrize("case,expected", [
    ("zero_diagonal", "invalid_input"),
    ("rank_zero_pivot", "rank_deficient"),
    ("scaling_overflow", "invalid_input"),
    ("beta_zero", "no_growth"),
    ("alpha_zero", "iteration_limit"),
])
def test_spai_failure_statuses(ctx, case, expected):
    """Failure fixtures of the reference (tests/release_cpu_snapshot.cpp:362-376,
    tests/test_spai.cpp:200-239)."""
    import bspai_hip

    P = _problems()

    def dense(m):
        m = np.asarray(m, dtype=float)
        r, c = np.nonzero(m)
        return P._from_coo(m.shape[0], r, c, m[r, c])

    tiny = [[4, 1, 0], [2, 5, 1], [0, 1, 3]]
    mats = {"zero_diagonal": ([[0, 1], [1, 1]], 1e-12, 3, 8),
            "rank_zero_pivot": ([[1, 2], [1, 2]], 1e-12, 3, 8),
            "scaling_overflow": ([[1e-320, 0], [0, 1]], 1e-12, 2, 8),
            "beta_zero": (tiny, 1e-12, 2, 0),
            "alpha_zero": (tiny, 1e-12, 0, 8)}
    m, eps, alpha, beta = mats[case]
    s = bspai_hip.construct_spai(bspai_hip.Matrix(dense(m), ctx), epsilon=eps, alpha=alpha,
                                 beta=beta)
    assert s.status == expected


def test_score_device_boundaries(ctx):
    """Restricted score known answers (reference tests/test_spai.cpp:240-250)."""
    import bspai_hip

    st, sc, cl = bspai_hip.evaluate_spai_score(4, 1, 2, ctx)
    assert st == "success" and sc == math.sqrt(3.5) and not cl
    st, sc, cl = bspai_hip.evaluate_spai_score(1, math.nextafter(1.0, 2.0), 1, ctx)
    assert st == "success" and cl and sc == 0
    bd = 1 + 16 * 2.0 ** -52
    assert bspai_hip.evaluate_spai_score(1, bd, 1, ctx)[0] == "success"
    assert bspai_hip.evaluate_spai_score(1, math.nextafter(bd, 2.0), 1, ctx)[0] == "invalid_score"
    assert bspai_hip.evaluate_spai_score(1, 1, 0, ctx)[0] == "zero_denominator"


def test_gmres_c1_vs_oracle(ctx):
    """One FP64 GMRES solve with the oracle's P: same status, iterations within 1."""
    import bspai_hip
    from oracle import pyoracle as O

    P = _problems()
    a = P.convection_diffusion_2d()
    s = O.construct_spai(a)
    p = P.Csr(a.n, s.p_ptr, s.p_col, s.p_val)
    b = P.equal_rhs(a.n)
    ref = O.gmres(a, p, b, 1e-8, 0)
    g = bspai_hip.left_preconditioned_gmres(bspai_hip.Matrix(a, ctx), bspai_hip.Matrix(p, ctx),
                                            b, 1e-8, 0)
    assert g.status == ref.status == "success"
    assert abs(g.iterations - ref.iterations) <= 1
    assert g.final_preconditioned_relative <= 1e-8
    x = np.asarray(g.solution)
    assert np.linalg.norm(x - ref.solution) <= 1e-6 * np.linalg.norm(ref.solution)


def test_ir_c1_fp64_vs_reference(ctx, c1_golden):
    """Full FP64 SPAI-GMRES-IR on the 64x64 system against cpu_fp64_ir's result."""
    import bspai_hip

    P = _problems()
    a = P.convection_diffusion_2d()
    b = P.equal_rhs(a.n)
    r = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(a, ctx), b, epsilon=0.3, beta=8,
                                tolerance=1e-8, max_iterations=0, max_outer_steps=20)
    assert r.status == str(c1_golden["ir_status"]) == "success"
    assert r.outer_steps == int(c1_golden["ir_outer_steps"])
    ref_it = c1_golden["ir_inner_iterations"]
    assert len(r.inner_iterations) == len(ref_it)
    assert all(abs(x - y) <= 1 for x, y in zip(r.inner_iterations, ref_it))
    assert r.normwise_backward_error <= 10 * 2.0 ** -53
    # x0 = P b and its diagnostics are bit-identical (same P, sequential SpMV):
    # ||r||_inf, normwise and componentwise backward error.  ||r||_2 is a
    # scaled sum of squares on the GPU (a hypot chain on the CPU): 1e-14.
    h0, g0 = r.history[0], c1_golden["ir_history"][0]
    assert np.array_equal(bits(h0[[0, 2, 3]]), bits(g0[[0, 2, 3]]))
    assert abs(h0[1] - g0[1]) <= 1e-14 * g0[1]
    x = np.asarray(r.solution)
    xr = c1_golden["ir_solution"]
    assert np.max(np.abs(x - xr)) <= 1e-10 * np.max(np.abs(xr))
    assert np.linalg.norm(b - a.multiply(x)) / np.linalg.norm(b) <= 1e-12


def test_ir_c1_mixed(ctx, c1_golden):
    """Mixed FP32/FP64: FP32 SPAI within 1e-5 of the FP64 P, IR to rel-res 1e-12."""
    import bspai_hip

    P = _problems()
    a = P.convection_diffusion_2d()
    b = P.equal_rhs(a.n)
    m = bspai_hip.Matrix(a, ctx)
    s32 = bspai_hip.construct_spai(m, epsilon=0.3, alpha=-(-a.n // 8), beta=8, precision="mixed")
    assert s32.status == "success"
    ptr, col, val = s32.preconditioner()
    gp, gc, gv = c1_golden["p_ptr"], c1_golden["p_col"], c1_golden["p_val"]
    diff_rows = 0
    rel = []
    for r in range(a.n):
        c32, v32 = col[ptr[r]:ptr[r + 1]], val[ptr[r]:ptr[r + 1]]
        c64, v64 = gc[gp[r]:gp[r + 1]], gv[gp[r]:gp[r + 1]]
        if not np.array_equal(c32, c64):
            diff_rows += 1
            continue
        # column k of M (= row k of P), normwise relative error
        rel.append(np.max(np.abs(v32 - v64)) / np.max(np.abs(v64)))
    rel = np.array(rel)
    print(f"mixed SPAI: pattern-differing columns {diff_rows}/{a.n}, max column rel "
          f"{rel.max():.3e}")
    assert diff_rows <= a.n // 100
    assert rel.max() <= 1e-5
    r = bspai_hip.spai_gmres_ir(m, b, precision="mixed", tolerance=1e-4, relres_target=1e-12)
    assert r.status == "success", r.message
    x = np.asarray(r.solution)
    assert np.linalg.norm(b - a.multiply(x)) / np.linalg.norm(b) <= 1e-12
    assert r.normwise_backward_error <= 10 * 2.0 ** -53


def test_batched_systems_match_single(ctx):
    """Block-diagonal batch: each system's result equals its own solve."""
    import bspai_hip

    P = _problems()
    rng = np.random.default_rng(3)
    blocks = []
    for side in (32, 40, 48):
        kappa = np.exp(rng.standard_normal((side, side)))
        blocks.append(P.variable_coefficient_2d(side, kappa, rng.uniform(0, 10)))
    big = P.block_diagonal(blocks)
    offs = np.cumsum([0] + [b.n for b in blocks])
    rhs = P.equal_rhs_blocks(offs)
    res = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(big, ctx, systems=offs), rhs)
    assert len(res) == 3
    x = np.asarray(res[0].solution)
    for s, blk in enumerate(blocks):
        single = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(blk, ctx), rhs[offs[s]:offs[s + 1]])
        assert res[s].status == single.status == "success"
        assert all(abs(p - q) <= 1 for p, q in zip(res[s].inner_iterations,
                                                   single.inner_iterations))
        xs = np.asarray(single.solution)
        xb = x[offs[s]:offs[s + 1]]
        # P's lane count may differ between batch and single (average row
        # length), which changes rounding only.
        assert np.max(np.abs(xb - xs)) <= 1e-12 * np.max(np.abs(xs))


def _seq_multiply(a, x):
    """y = A x with sequential per-row sums (the reference's order)."""
    y = np.zeros(a.n)
    lens = np.diff(a.ptr)
    for k in range(int(lens.max())):
        rows = np.nonzero(lens > k)[0]
        q = a.ptr[rows] + k
        y[rows] = y[rows] + a.val[q] * x[a.col[q]]
    return y


def test_many_systems_batched_equals_looped(ctx):
    """A slice of the many-systems workload (24 variable-coefficient
    diffusion-convection systems, L in 32..96, log-normal conductivity): one
    batched SPAI + IR launch sequence returns, per system, exactly what the
    single-system plain path returns (same statuses and iterations, same
    solution bits: every kernel reduces within a system in the same order)."""
    import bspai_hip

    P = _problems()
    sb = P.many_systems(count=24, lmax=96)
    mb = bspai_hip.Matrix(sb.matrix, ctx, systems=sb.offsets)
    kw = dict(tolerance=1e-8, max_iterations=400)
    res = bspai_hip.spai_gmres_ir(mb, sb.rhs, **kw)
    x = np.asarray(res[0].solution)
    for s in range(len(sb.offsets) - 1):
        o0, o1 = int(sb.offsets[s]), int(sb.offsets[s + 1])
        single = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(sb.block(s), ctx), sb.rhs[o0:o1], **kw)
        assert res[s].status == single.status == "success"
        assert res[s].inner_iterations == single.inner_iterations
        assert np.array_equal(bits(x[o0:o1]), bits(np.asarray(single.solution)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spmv_autotune_keeps_bits(monkeypatch, precision):
    """The SpMV autotuner picks among order-preserving kernels only: an IR
    solve with tuning on returns the same bits and iteration counts as with
    tuning off (fresh contexts, 32^3 convection-diffusion)."""
    import bspai_hip

    P = _problems()
    a = P.convection_diffusion_3d(32)
    b = P.equal_rhs(a.n)
    kw = dict(precision=precision, tolerance=1e-8 if precision == "fp64" else 2e-5,
              max_iterations=200, relres_target=1e-12)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("BSPAI_SPMV_AUTOTUNE", mode)
        c = bspai_hip.Context(0)
        r = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(a, c), b, **kw)
        r2 = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(a, c), b, **kw)  # cached choice
        assert r.status == r2.status == "success"
        assert r.inner_iterations == r2.inner_iterations
        assert np.array_equal(bits(np.asarray(r.solution)), bits(np.asarray(r2.solution)))
        out[mode] = r
    assert out["0"].inner_iterations == out["1"].inner_iterations
    assert np.array_equal(bits(np.asarray(out["0"].solution)), bits(np.asarray(out["1"].solution)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spmv_col16_keeps_bits(monkeypatch, precision):
    """The streaming SpMV on 16-bit column offsets (2 bytes per entry instead
    of 4) gives the same columns and sums: IR solves with it forced and with
    32-bit columns return the same bits and iteration counts (32^3
    convection-diffusion, whose offsets to a 256-row block fit 16 bits).  A
    matrix with a coupling 40,000 columns away cannot use the offsets, and
    keeps 32-bit columns and the same bits."""
    import bspai_hip

    P = _problems()
    n = 50000
    i = np.arange(n)
    far = P._from_coo(n, np.concatenate([i, i[1:], i[:-1], i]),
                      np.concatenate([i, i[:-1], i[1:], (i + 40000) % n]),
                      np.concatenate([np.full(n, 4.0), np.full(n - 1, -1.0),
                                      np.full(n - 1, -1.2), np.full(n, 0.5)]))
    for a, kw in ((P.convection_diffusion_3d(32),
                   dict(tolerance=1e-8 if precision == "fp64" else 2e-5, max_iterations=200,
                        relres_target=1e-12)),
                  (far, dict(tolerance=1e-8 if precision == "fp64" else 1e-4,
                             max_iterations=100, relres_target=1e-12))):
        b = P.equal_rhs(a.n)
        out = {}
        for mode in ("0", "1"):
            monkeypatch.setenv("BSPAI_SPMV_C16", mode)
            c = bspai_hip.Context(0)
            out[mode] = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(a, c), b, precision=precision,
                                                **kw)
            assert out[mode].status == "success"
        assert out["0"].inner_iterations == out["1"].inner_iterations
        assert np.array_equal(bits(np.asarray(out["0"].solution)),
                              bits(np.asarray(out["1"].solution)))


@pytest.mark.parametrize("case", ["c1_fp64", "c1_mixed_floor", "cube32_mixed", "c1_mixed_cap"])
def test_deferred_verification_same_as_every_step(monkeypatch, case):
    """A single system verifies triggered steps at the host's check points
    (the held step is verified, then the loop resumes after it).  Against
    per-step verification: same statuses, inner iteration counts, outer steps
    and solution bits -- including mixed solves whose tolerance sits at the
    FP32 floor, where triggered steps fail verification and the loop resumes,
    and a cap that ends the inner solves."""
    import bspai_hip

    P = _problems()
    a = P.convection_diffusion_3d(32) if case == "cube32_mixed" else P.convection_diffusion_2d()
    b = P.equal_rhs(a.n)
    kw = {"c1_fp64": dict(precision="fp64", tolerance=1e-8),
          "c1_mixed_floor": dict(precision="mixed", tolerance=2e-7, max_iterations=300),
          "cube32_mixed": dict(precision="mixed", tolerance=2e-5),
          "c1_mixed_cap": dict(precision="mixed", tolerance=1e-6, max_iterations=37,
                               max_outer_steps=4)}[case]
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("BSPAI_GMRES_VERIFY_EVERY_STEP", mode)
        c = bspai_hip.Context(0)
        out[mode] = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(a, c), b, **kw)
    d, e = out["0"], out["1"]
    print(case, d.status, d.outer_steps, d.inner_iterations, d.inner_statuses)
    assert d.status == e.status and d.outer_steps == e.outer_steps
    assert d.inner_iterations == e.inner_iterations
    assert d.inner_statuses == e.inner_statuses
    assert np.array_equal(bits(np.asarray(d.solution)), bits(np.asarray(e.solution)))


def test_gmres_second_inner_solve_same_input(ctx):
    """The second inner solve of the 64x64 IR, given the oracle's own residual
    r1 = b - A x1: GPU and oracle iteration counts agree within 1."""
    import bspai_hip
    from oracle import pyoracle as O

    P = _problems()
    a = P.convection_diffusion_2d()
    b = P.equal_rhs(a.n)
    s = O.construct_spai(a)
    p = P.Csr(a.n, s.p_ptr, s.p_col, s.p_val)
    one = O.spai_gmres_ir(a, b, p=p, max_outer_steps=1)
    r1 = b - _seq_multiply(a, one.solution)
    ref = O.gmres(a, p, r1, 1e-8, 0)
    g = bspai_hip.left_preconditioned_gmres(bspai_hip.Matrix(a, ctx), bspai_hip.Matrix(p, ctx),
                                            r1, 1e-8, 0)
    assert g.status == ref.status == "success"
    assert abs(g.iterations - ref.iterations) <= 1
    assert np.linalg.norm(np.asarray(g.solution) - ref.solution) <= 1e-6 * np.linalg.norm(
        ref.solution)


@pytest.mark.parametrize("case", ["c1_fp64", "c1_mixed", "cube32_mixed", "batched"])
def test_graph_replay_same_bits(monkeypatch, case):
    """Arnoldi steps replayed from captured graphs (the step read from device
    memory, launch shapes sized for the cap) against the same kernels
    launched one by one: identical statuses, counts and solution bits."""
    import bspai_hip

    P = _problems()
    if case == "batched":
        sb = P.many_systems(count=12, lmax=80)
        a, b, systems = sb.matrix, sb.rhs, sb.offsets
    else:
        a = P.convection_diffusion_3d(32) if case.startswith("cube") else P.convection_diffusion_2d()
        b, systems = P.equal_rhs(a.n), None
    prec = "fp64" if case.endswith("fp64") or case == "batched" else "mixed"
    kw = dict(precision=prec, tolerance=1e-8 if prec == "fp64" else 1e-4, max_iterations=200,
              relres_target=1e-12 if prec == "mixed" else 0.0)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("BSPAI_GMRES_GRAPH", mode)
        ctx = bspai_hip.Context(0)
        m = bspai_hip.Matrix(a, ctx, systems=systems)
        r = bspai_hip.spai_gmres_ir(m, b, **kw)
        rs = r if isinstance(r, list) else [r]
        out[mode] = ([(x.status, x.outer_steps, x.inner_iterations) for x in rs],
                     bits(np.asarray(rs[0].solution)), ctx.graph_stats())
    assert out["0"][0] == out["1"][0]
    assert all(s == "success" for s, _, _ in out["1"][0])
    assert np.array_equal(out["0"][1], out["1"][1])
    assert out["0"][2] == (0, 0)
    assert out["1"][2][1] > 0  # graphs were replayed


@pytest.mark.parametrize("case", ["c1_fp64", "cube32_mixed", "long_fp64", "batched"])
def test_gram_solve_in_update_same_bits(monkeypatch, case):
    """The Gram solve inside the update pass (S = 4 for launches up to step
    255, 8 up to 511) against the separate one-wave solve kernel at every
    step: identical statuses, counts and solution bits.  The long case (a
    2-D system of 288^2 rows with the first fit of every SPAI column only,
    epsilon 2, to 1e-10) runs past step 511, where both take the separate
    kernel."""
    import bspai_hip

    P = _problems()
    if case == "batched":
        sb = P.many_systems(count=12, lmax=80)
        a, b, systems = sb.matrix, sb.rhs, sb.offsets
    elif case.startswith("long"):
        a = P.convection_diffusion_2d(288)
        b, systems = P.equal_rhs(a.n), None
    else:
        a = P.convection_diffusion_3d(32) if case.startswith("cube") else P.convection_diffusion_2d()
        b, systems = P.equal_rhs(a.n), None
    prec = "mixed" if case.endswith("mixed") else "fp64"
    kw = dict(precision=prec,
              tolerance=1e-10 if case.startswith("long") else 1e-8 if prec == "fp64" else 1e-4,
              max_iterations=1500 if case.startswith("long") else 200,
              relres_target=1e-12 if prec == "mixed" else 0.0,
              epsilon=2.0 if case.startswith("long") else 0.3)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BSPAI_GRAM_FUSED", mode)
        ctx = bspai_hip.Context(0)
        m = bspai_hip.Matrix(a, ctx, systems=systems)
        r = bspai_hip.spai_gmres_ir(m, b, **kw)
        rs = r if isinstance(r, list) else [r]
        out[mode] = ([(x.status, x.outer_steps, x.inner_iterations) for x in rs],
                     [bits(np.asarray(x.solution)) for x in rs])
    assert out["0"][0] == out["1"][0]
    assert all(s == "success" for s, _, _ in out["1"][0])
    for x0, x1 in zip(out["0"][1], out["1"][1]):
        assert np.array_equal(x0, x1)
    if case.startswith("long"):
        assert max(out["1"][0][0][2]) > 511


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spai_bucket_autotune_keeps_bits(monkeypatch, precision):
    """The capacity-bucket autotuner times launch shapes of the workgroup SPAI
    kernel on a slice of a bucket's worklist and fits the rest with the
    fastest; a column's fit does not depend on the shape.  Power law
    n = 5,000 with the tuning threshold lowered so that it tunes: patterns,
    coefficients and per-column records equal the untuned build bit for bit,
    and a second build of the same shape reuses the cached shape."""
    import bspai_hip

    P = _problems()
    a = P.power_law(n=5000)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("BSPAI_SPAI_AUTOTUNE", mode)
        monkeypatch.setenv("BSPAI_SPAI_TUNE_MIN", "64")
        c = bspai_hip.Context(0)
        m = bspai_hip.Matrix(a, c)
        c.set_timing("spai")
        s = bspai_hip.construct_spai(m, 0.2, -(-a.n // 8), 8, precision=precision)
        labels = list(c.kernel_times())
        s2 = bspai_hip.construct_spai(m, 0.2, -(-a.n // 8), 8, precision=precision)
        labels2 = list(c.kernel_times())
        c.set_timing(False)
        assert s.status == s2.status == "success" and s.passes >= 2
        if mode == "1":
            assert any(k.endswith(", tuned)") for k in labels), labels
            assert not any(k.endswith(", tuned)") for k in labels2[len(labels):]), labels2
        out[mode] = (s, s2)
    ref = out["0"][0]
    for s in (out["1"][0], out["1"][1], out["0"][1]):
        ptr, col, val = s.preconditioner()
        rp, rc, rv = ref.preconditioner()
        assert np.array_equal(ptr, rp) and np.array_equal(col, rc)
        assert np.array_equal(np.asarray(val).view(np.uint64), np.asarray(rv).view(np.uint64))
        for k, v in s.columns().items():
            assert np.array_equal(np.asarray(v), np.asarray(ref.columns()[k])), k
