"""The sliced SpMV (one wavefront per 64-row slice, entries column-major per
slice) against the sequential CSR row sum of the reference's
CsrMatrix::multiply (src/cpu.cpp:80-94): bit-exact in FP64 and, for the FP32
product, bit-exact against the same sequential order in binary32.  Shapes
cover slices cut by the end of the matrix and by system boundaries (a
batched layout), empty rows, long rows, and the layout's two refusals."""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def _seq_product(a, x, dtype):
    """Row sums in CSR order, every operation rounded to `dtype`."""
    val = a.val.astype(dtype)
    x = np.asarray(x, dtype)
    y = np.zeros(a.n, dtype)
    for r in range(a.n):
        s = dtype(0)
        for q in range(int(a.ptr[r]), int(a.ptr[r + 1])):
            s = dtype(s + dtype(val[q] * x[a.col[q]]))
        y[r] = s
    return y


def _random_uneven(n, seed, maxlen=40, empty_every=17):
    """Rows of 0..maxlen entries at random columns (some rows empty)."""
    from bspai_hip import problems as P

    rng = np.random.default_rng(seed)
    lens = rng.integers(0, maxlen + 1, size=n)
    lens[::empty_every] = 0
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(0, n, size=rows.size)
    key = np.unique(rows * n + cols)
    rows, cols = key // n, key % n
    vals = rng.standard_normal(rows.size)
    return P._from_coo(n, rows, cols, vals)


def _cases():
    from bspai_hip import problems as P

    batch = P.many_systems(count=12, lmin=8, lmax=40)
    return {
        "c1": (P.convection_diffusion_2d(), None),
        "uneven3000": (_random_uneven(3000, 11), None),
        "batch12": (batch.matrix, batch.offsets),
    }


@pytest.mark.parametrize("case", ["c1", "uneven3000", "batch12"])
@pytest.mark.parametrize("kind", ["sliced", "sliced16"])
def test_sliced_bit_exact(ctx, case, kind):
    import bspai_hip
    from bspai_hip import _lib

    a, offsets = _cases()[case]
    lanes = _lib.SPMV_SLICED if kind == "sliced" else _lib.SPMV_SLICED16
    m = bspai_hip.Matrix(a, ctx, systems=offsets)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(a.n)
    y = m.multiply(x, lanes=lanes)
    assert np.array_equal(bits(y), bits(_seq_product(a, x, np.float64)))
    x32 = x.astype(np.float32)
    y32 = m.multiply_f32(x32, lanes=lanes)
    ref32 = _seq_product(a, x32, np.float32)
    assert np.array_equal(y32.view(np.uint32), ref32.view(np.uint32))


def test_sliced_values_follow_set_values(ctx):
    """New values reach the sliced copy (its values are refilled)."""
    import bspai_hip
    from bspai_hip import _lib

    from bspai_hip import problems as P

    a = P.convection_diffusion_2d()
    m = bspai_hip.Matrix(a, ctx)
    x = np.random.default_rng(5).standard_normal(a.n)
    m.multiply(x, lanes=_lib.SPMV_SLICED)
    m.multiply_f32(x.astype(np.float32), lanes=_lib.SPMV_SLICED)
    a2 = P.Csr(a.n, a.ptr, a.col, a.val * 3.0 - 1.0)
    m.set_values(a2.val)
    assert np.array_equal(bits(m.multiply(x, lanes=_lib.SPMV_SLICED)),
                          bits(_seq_product(a2, x, np.float64)))
    y32 = m.multiply_f32(x.astype(np.float32), lanes=_lib.SPMV_SLICED)
    assert np.array_equal(y32.view(np.uint32),
                          _seq_product(a2, x.astype(np.float32), np.float32).view(np.uint32))


def test_sliced_refusals(ctx):
    """A row of more than 255 entries has no sliced layout; a column offset
    beyond int16 has only the int32 form."""
    import bspai_hip
    from bspai_hip import _lib

    from bspai_hip import problems as P

    n = 600
    rows = np.concatenate([np.zeros(299, np.int64), np.arange(n)])
    cols = np.concatenate([1 + 2 * np.arange(299), np.arange(n)])
    long_row = P._from_coo(n, rows, cols, np.ones(rows.size))
    m = bspai_hip.Matrix(long_row, ctx)
    x = np.ones(n)
    with pytest.raises(ValueError, match="255"):
        m.multiply(x, lanes=_lib.SPMV_SLICED)
    assert np.array_equal(m.multiply(x, lanes=1), _seq_product(long_row, x, np.float64))

    n = 40000
    rows = np.concatenate([np.arange(n), [0, n - 1]])
    cols = np.concatenate([np.arange(n), [n - 1, 0]])
    far = P._from_coo(n, rows, cols, np.linspace(1.0, 2.0, rows.size))
    m = bspai_hip.Matrix(far, ctx)
    x = np.random.default_rng(9).standard_normal(n)
    with pytest.raises(ValueError, match="16 bits"):
        m.multiply(x, lanes=_lib.SPMV_SLICED16)
    assert np.array_equal(bits(m.multiply(x, lanes=_lib.SPMV_SLICED)),
                          bits(_seq_product(far, x, np.float64)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sliced_in_solve_keeps_bits(monkeypatch, precision):
    """IR solves with the sliced SpMV forced into every Arnoldi step (replayed
    graphs included) return the same bits and iteration counts as with the
    streaming kernel (32^3 convection-diffusion; batched 12-system layout)."""
    import bspai_hip
    from bspai_hip import problems as P

    a = P.convection_diffusion_3d(32)
    b = P.equal_rhs(a.n)
    batch = P.many_systems(count=12, lmin=8, lmax=40)
    kw = dict(precision=precision, tolerance=1e-8 if precision == "fp64" else 2.0 ** -16,
              max_iterations=200, relres_target=1e-12)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("BSPAI_SPMV_AUTOTUNE", "1")
        monkeypatch.setenv("BSPAI_SPMV_SLICED", mode)
        c = bspai_hip.Context(0)
        r = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(a, c), b, **kw)
        rb = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(batch.matrix, c, systems=batch.offsets),
                                     batch.rhs, **dict(kw, max_iterations=0))
        assert r.status == "success"
        out[mode] = (r, rb)
    (r0, rb0), (r1, rb1) = out["0"], out["1"]
    assert r0.inner_iterations == r1.inner_iterations
    assert np.array_equal(bits(np.asarray(r0.solution)), bits(np.asarray(r1.solution)))
    for s0, s1 in zip(rb0, rb1):
        assert s0.status == s1.status and s0.inner_iterations == s1.inner_iterations
        assert np.array_equal(bits(np.asarray(s0.solution)), bits(np.asarray(s1.solution)))


def _dominant(n, lens, seed):
    """Diagonally dominant rows with lens[r] off-diagonal entries."""
    from bspai_hip import problems as P

    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for r, k in enumerate(lens):
        c = rng.choice(n - 1, size=int(k), replace=False)
        cols.append(c + (c >= r))
        rows.append(np.full(int(k), r))
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    vals = rng.uniform(-1.0, 1.0, size=rows.size)
    diag = np.zeros(n)
    np.add.at(diag, rows, np.abs(vals))
    return P._from_coo(n, np.concatenate([rows, np.arange(n)]),
                       np.concatenate([cols, np.arange(n)]), np.concatenate([vals, diag + 1.0]))


def test_tuned_choice_of_a_shape_fits_each_matrix(monkeypatch):
    """The SpMV choice is tuned per shape (n, nnz); a second matrix of the same
    shape that cannot take the sliced layout (a row of 300 entries) still
    solves, with the bits of a solve in a fresh context."""
    import bspai_hip
    from bspai_hip import problems as P

    monkeypatch.setenv("BSPAI_SPMV_AUTOTUNE", "1")
    monkeypatch.delenv("BSPAI_SPMV_SLICED", raising=False)
    n = 3000
    lens_b = np.full(n, 6)
    lens_b[7] = 300
    lens_a = np.full(n, 6)
    lens_a[: (300 - 6)] += 1  # the same total count of entries
    a, b_mat = _dominant(n, lens_a, 1), _dominant(n, lens_b, 2)
    assert a.nnz == b_mat.nnz
    rhs = P.equal_rhs(n)
    kw = dict(precision="fp64", tolerance=1e-8, max_iterations=200)
    c = bspai_hip.Context(0)
    ra = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(a, c), rhs, **kw)
    rb = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(b_mat, c), rhs, **kw)
    fresh = bspai_hip.spai_gmres_ir(bspai_hip.Matrix(b_mat, bspai_hip.Context(0)), rhs, **kw)
    assert ra.status == rb.status == fresh.status == "success"
    assert rb.inner_iterations == fresh.inner_iterations
    assert np.array_equal(bits(np.asarray(rb.solution)), bits(np.asarray(fresh.solution)))
