"""The benchmark configurations under test against the oracle, at full size
where the oracle allows it and by size-independent properties beyond.

  * 3-D 128^3 convection-diffusion (n = 2,097,152, the benchmark system): the
    FP64 SPAI bit for bit against the oracle's column subset on 4,096 columns
    drawn from every boundary class (corners, edges, faces, interior): pattern,
    fits, growths and coefficient bits; both the plain FP64 and the mixed path
    reach ||b-Ax||/||b|| <= 1e-12 with eta <= 10u; the FP32 SPAI against the
    FP64 one, pattern differences counted, M within 1e-5 per column.
  * 3-D 32^3 (n = 32,768), a size the full oracle covers: the whole FP64 SPAI
    bit for bit, IR status equal and inner iteration counts within +-1.
  * The 256-system many-systems batch (L in 32..256, 6.5M rows): batched SPAI
    + IR equal to the looped single-system path bit for bit, and three of its
    systems (smallest, largest, one more) against the oracle.
Reference bars: tests/test_cuda_spai.cpp:44-53 (per-column status, pattern,
fits and growths equal to the CPU), tests/test_cuda_gmres.cpp:84-90 (status
and iteration count equal to the CPU, solution within 1e-12).
"""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def _P():
    from bspai_hip import problems

    return problems


def _O():
    from oracle import pyoracle

    return pyoracle


@pytest.fixture(scope="module")
def cube128(ctx):
    import bspai_hip

    a = _P().convection_diffusion_3d(128)
    return a, bspai_hip.Matrix(a, ctx)


def _classes(side, rng):
    """4,096 columns: the 8 corners, 256 edge, 1,024 face and 2,808 interior
    points of a side^3 grid (x fastest)."""
    idx = np.arange(side ** 3)
    x, y, z = idx % side, (idx // side) % side, idx // (side * side)
    nb = sum(((c == 0) | (c == side - 1)).astype(int) for c in (x, y, z))
    pick = [idx[nb == 3]]
    for cls, count in ((2, 256), (1, 1024), (0, 2808)):
        pick.append(rng.choice(idx[nb == cls], count, replace=False))
    return np.sort(np.concatenate(pick))


def test_c2_fp64_spai_columns_bit_exact(cube128):
    import bspai_hip

    a, m = cube128
    s = bspai_hip.construct_spai(m, 0.3, -(-a.n // 8), 8)
    assert s.status == "success"
    cols = _classes(128, np.random.default_rng(128))
    st, fi, gr, pt, ix, cf = _O().spai_subset(a, cols, epsilon=0.3)
    pp, pi, pc = s.patterns()
    info = s.columns()
    for i, k in enumerate(cols):
        assert list(pi[pp[k]:pp[k + 1]]) == list(ix[pt[i]:pt[i + 1]]), k
        assert np.array_equal(bits(pc[pp[k]:pp[k + 1]]), bits(cf[pt[i]:pt[i + 1]])), k
        assert info["fits"][k] == fi[i] and info["growths"][k] == gr[i], k
        assert info["status"][k] == st[i] == 0, k


def _row_compare(p32, p64):
    """Rows of P (= columns of M) compared: the count whose patterns differ,
    and the largest per-row normwise relative difference on the others."""
    ptr32, col32, val32 = p32
    ptr64, col64, val64 = p64
    n = len(ptr64) - 1
    l32, l64 = np.diff(ptr32), np.diff(ptr64)
    same = l32 == l64
    rows = np.nonzero(same)[0]
    L = l64[rows]
    offs = np.arange(L.sum()) - np.repeat(np.cumsum(L) - L, L)
    i32 = np.repeat(ptr32[:-1][rows], L) + offs
    i64 = np.repeat(ptr64[:-1][rows], L) + offs
    starts = np.cumsum(L) - L
    mism = np.add.reduceat((col32[i32] != col64[i64]).astype(np.int64), starts)
    eq = mism == 0
    num = np.maximum.reduceat(np.abs(val32[i32] - val64[i64]), starts)
    den = np.maximum.reduceat(np.abs(val64[i64]), starts)
    rel = (num / den)[eq]
    return int(n - eq.sum()), float(rel.max())


def test_c2_plain_and_mixed_reach_target(cube128):
    """Both precisions on the benchmark system: rel-res <= 1e-12, eta <= 10u;
    the FP32 preconditioner's pattern differences counted, M within 1e-5."""
    import bspai_hip

    a, m = cube128
    b = _P().equal_rhs(a.n)
    A = a.to_scipy()
    common = dict(epsilon=0.3, alpha=0, beta=8, max_iterations=200, max_outer_steps=20,
                  relres_target=1e-12)
    out = {}
    for prec, tol in (("fp64", 1e-8), ("mixed", 2.0 ** -16)):
        r = bspai_hip.spai_gmres_ir(m, b, precision=prec, tolerance=tol, **common)
        x = np.asarray(r.solution)
        relres = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        out[prec] = r
        print(f"{prec}: {r.status} outer {r.outer_steps} inner {r.inner_iterations} "
              f"relres {relres:.3e} eta {r.normwise_backward_error:.3e}")
        assert r.status == "success", r.message
        assert relres <= 1e-12
        assert r.normwise_backward_error <= 10 * 2.0 ** -53
    # FP64: the reference's method; inner counts as measured for this system
    assert all(abs(k - 157) <= 2 for k in out["fp64"].inner_iterations)
    s64 = bspai_hip.construct_spai(m, 0.3, -(-a.n // 8), 8)
    s32 = bspai_hip.construct_spai(m, 0.3, -(-a.n // 8), 8, precision="mixed")
    assert s64.status == s32.status == "success"
    ndiff, rel = _row_compare(s32.preconditioner(), s64.preconditioner())
    print(f"FP32 SPAI vs FP64: pattern-differing columns {ndiff}/{a.n}, max column rel {rel:.3e}")
    assert ndiff <= a.n // 100
    assert rel <= 1e-5


def test_cube32_full_oracle(ctx):
    import bspai_hip

    a = _P().convection_diffusion_3d(32)
    b = _P().equal_rhs(a.n)
    m = bspai_hip.Matrix(a, ctx)
    alpha = -(-a.n // 8)
    s = bspai_hip.construct_spai(m, 0.3, alpha, 8)
    ref = _O().construct_spai(a, 0.3, alpha, 8)
    assert s.status == ref.status == "success"
    ptr, col, val = s.preconditioner()
    assert np.array_equal(ptr, ref.p_ptr) and np.array_equal(col, ref.p_col)
    assert np.array_equal(bits(val), bits(ref.p_val))
    assert np.array_equal(s.columns()["fits"], ref.fits)
    assert np.array_equal(s.columns()["growths"], ref.growths)
    r = bspai_hip.spai_gmres_ir(m, b, epsilon=0.3, max_iterations=200)
    ir = _O().spai_gmres_ir(a, b, 0.3, max_iterations=200)
    assert r.status == ir.status == "success"
    assert r.outer_steps == ir.outer_steps
    assert all(abs(x - y) <= 1 for x, y in zip(r.inner_iterations, ir.inner_iterations))
    xr = np.asarray(ir.solution)
    assert np.max(np.abs(np.asarray(r.solution) - xr)) <= 1e-10 * np.max(np.abs(xr))


@pytest.fixture(scope="module")
def batch256(ctx):
    import bspai_hip

    sb = _P().many_systems()
    return sb, bspai_hip.Matrix(sb.matrix, ctx, systems=sb.offsets)


def test_c4_full_batch_equals_looped(ctx, batch256):
    """All 256 systems: one batched solve against 256 single solves, bits equal."""
    import bspai_hip

    sb, mb = batch256
    kw = dict(tolerance=1e-8, max_iterations=400)
    res = bspai_hip.spai_gmres_ir(mb, sb.rhs, **kw)
    x = np.asarray(res[0].solution)
    ns = len(sb.offsets) - 1
    assert ns == 256
    for s in range(ns):
        o0, o1 = int(sb.offsets[s]), int(sb.offsets[s + 1])
        single = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(sb.block(s), ctx), sb.rhs[o0:o1], **kw)
        assert res[s].status == single.status == "success", s
        assert res[s].inner_iterations == single.inner_iterations, s
        assert np.array_equal(bits(x[o0:o1]), bits(np.asarray(single.solution))), s


def test_c4_systems_vs_oracle(ctx, batch256):
    """Smallest, largest and a middle system of the batch against the oracle:
    SPAI bits, IR status, outer steps and inner counts within +-1."""
    import bspai_hip

    sb, mb = batch256
    sizes = np.diff(sb.offsets)
    picks = [int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argsort(sizes)[len(sizes) // 2])]
    res = bspai_hip.spai_gmres_ir(mb, sb.rhs, tolerance=1e-8, max_iterations=400)
    x = np.asarray(res[0].solution)
    for s in picks:
        o0, o1 = int(sb.offsets[s]), int(sb.offsets[s + 1])
        blk = sb.block(s)
        sp = bspai_hip.construct_spai(bspai_hip.Matrix(blk, ctx), 0.3, -(-blk.n // 8), 8)
        ref = _O().construct_spai(blk, 0.3, -(-blk.n // 8), 8)
        assert sp.status == ref.status == "success"
        ptr, col, val = sp.preconditioner()
        assert np.array_equal(ptr, ref.p_ptr) and np.array_equal(col, ref.p_col)
        assert np.array_equal(bits(val), bits(ref.p_val))
        ir = _O().spai_gmres_ir(blk, sb.rhs[o0:o1], 0.3, max_iterations=400)
        assert res[s].status == ir.status == "success"
        assert res[s].outer_steps == ir.outer_steps
        assert all(abs(p - q) <= 1 for p, q in zip(res[s].inner_iterations, ir.inner_iterations))
        xr = np.asarray(ir.solution)
        assert np.max(np.abs(x[o0:o1] - xr)) <= 1e-10 * np.max(np.abs(xr))


def test_batched_unequal_sizes_own_caps(ctx):
    """Systems of unequal size in one batch, cap 0 and an unreachable
    tolerance: each system stops at its own dimension with the single-system
    status (nonconvergence at the full dimension), as the oracle's solve of
    that system alone does."""
    import bspai_hip

    P = _P()
    rng = np.random.default_rng(5)
    blocks = []
    for n in (5, 12, 30):
        d = rng.standard_normal((n, n)) + 6 * np.eye(n)
        d[np.abs(d) < 0.8] = 0.0
        np.fill_diagonal(d, 6 + rng.random(n))
        r, c = np.nonzero(d)
        blocks.append(P._from_coo(n, r, c, d[r, c]))
    big = P.block_diagonal(blocks)
    offs = np.cumsum([0] + [b.n for b in blocks])
    ident = P._from_coo(big.n, np.arange(big.n), np.arange(big.n), np.ones(big.n))
    rhs = rng.standard_normal(big.n)
    mb = bspai_hip.Matrix(big, ctx, systems=offs)
    pb = bspai_hip.Matrix(ident, ctx, systems=offs)
    out = bspai_hip.left_preconditioned_gmres(mb, pb, rhs, tolerance=1e-300, max_iterations=0)
    for s, blk in enumerate(blocks):
        o0, o1 = offs[s], offs[s + 1]
        pid = P._from_coo(blk.n, np.arange(blk.n), np.arange(blk.n), np.ones(blk.n))
        ref = _O().gmres(blk, pid, rhs[o0:o1], 1e-300, 0)
        assert out[s].status == ref.status == "nonconvergence", (s, out[s].status, ref.status)
        assert out[s].iterations == ref.iterations == blk.n
    # a cap above the smallest system is rejected, for GMRES and IR alike
    with pytest.raises(ValueError):
        bspai_hip.left_preconditioned_gmres(mb, pb, rhs, max_iterations=6)
    with pytest.raises(ValueError):
        bspai_hip.refine_with_preconditioner(mb, pb, rhs, max_iterations=6)


def test_torch_inputs_of_other_dtypes(ctx):
    """Device tensors in other dtypes (int64 structure, float32 values and
    vectors) are converted, not read as raw int32 / float64 memory: the result
    equals the FP64 path on the same (exactly widened) values."""
    import torch

    import bspai_hip

    P = _P()
    a = P.convection_diffusion_2d(32)
    v32 = a.val.astype(np.float32)
    dev = torch.device("cuda:0")
    m_t = bspai_hip.Matrix((torch.tensor(a.ptr, dtype=torch.int64, device=dev),
                            torch.tensor(a.col, dtype=torch.int64, device=dev),
                            torch.tensor(v32, device=dev)), ctx)
    m_h = bspai_hip.Matrix((a.ptr, a.col, v32.astype(np.float64)), ctx)
    x = np.random.default_rng(1).standard_normal(a.n).astype(np.float32)
    y_t = m_t.multiply(torch.tensor(x, device=dev)).cpu().numpy()
    y_h = m_h.multiply(x.astype(np.float64))
    assert np.array_equal(bits(y_t), bits(y_h))
    b = P.equal_rhs(a.n)
    r_t = bspai_hip.spai_gmres_ir(m_t, torch.tensor(b, dtype=torch.float32, device=dev))
    r_h = bspai_hip.spai_gmres_ir(m_h, b.astype(np.float32).astype(np.float64))
    assert r_t.status == r_h.status == "success"
    assert np.array_equal(bits(r_t.solution.cpu().numpy()), bits(np.asarray(r_h.solution)))
    with pytest.raises(ValueError):
        m_h.multiply(np.ones(a.n + 1))
    with pytest.raises(ValueError):
        bspai_hip.spai_gmres_ir(m_h, torch.ones(a.n - 1, dtype=torch.float64, device=dev))


@pytest.mark.parametrize("where", ["host", "device"])
def test_matrix_validation(ctx, where):
    """Non-canonical CSR is rejected on both paths (the kernels index by it):
    unsorted or repeated columns, columns out of range, offsets out of order,
    and, for a batch, an entry outside its row's system."""
    import torch

    import bspai_hip

    def make(ptr, col, val, systems=None):
        if where == "device":
            dev = torch.device("cuda:0")
            ptr = torch.tensor(ptr, dtype=torch.int32, device=dev)
            col = torch.tensor(col, dtype=torch.int32, device=dev)
            val = torch.tensor(val, dtype=torch.float64, device=dev)
        return bspai_hip.Matrix((np.asarray(ptr) if where == "host" else ptr,
                                 np.asarray(col) if where == "host" else col,
                                 np.asarray(val, dtype=float) if where == "host" else val),
                                ctx, systems=systems)

    ok = make([0, 2, 4], [0, 1, 0, 1], [2.0, 1, 1, 2])
    assert ok.n == 2
    bad = [
        ([0, 2, 4], [1, 0, 0, 1], [2.0, 1, 1, 2], None, "strictly increasing"),
        ([0, 2, 4], [0, 0, 0, 1], [2.0, 1, 1, 2], None, "strictly increasing"),
        ([0, 2, 4], [0, 2, 0, 1], [2.0, 1, 1, 2], None, "out of range"),
        ([0, 5, 4], [0, 1, 0, 1], [2.0, 1, 1, 2], None, "row pointers"),
        ([0, 3, 2], [0, 1, 0, 1], [2.0, 1, 1, 2], None, "row pointers"),
        ([0, 2, 4], [0, 1, 0, 1], [2.0, 1, 1, float("nan")], None, "finite"),
        ([0, 2, 4], [0, 1, 0, 1], [2.0, 1, 1, 2], [0, 1, 2], "block-diagonal"),
    ]
    for ptr, col, val, systems, msg in bad:
        with pytest.raises(ValueError, match=msg):
            make(ptr, col, val, systems)
