"""Deterministic sparse test systems for the SPAI-GMRES-IR path.

Every generator returns a canonical CSR triple (row pointers, column indices,
values) with int64 indices: rows sorted, no duplicates, no stored zeros -- the
form `CsrMatrix::from_entries` produces (reference src/cpu.cpp:49-78).

Workloads:
  convection_diffusion_2d   5-point upwind stencil, 64x64 grid (n=4096)
  convection_diffusion_3d   7-point upwind stencil, 128^3 grid (n=2,097,152)
  power_law                 random nonsymmetric, power-law row lengths (n=500,000)
  many_systems              256 variable-coefficient 2-D systems, side 32..256
  diffusion_edges_2d        2-D edge-conductivity diffusion A(theta), 256^2
  band                      the reference's banded SPAI family
                            (reference tests/test_cuda_spai_sparse.cpp:7)

Stencil matrices are scaled by h^2 with unit diffusion, so their entries are
small integers and are exact in binary64 and in a Matrix Market file.  The
cell Peclet number of a direction is v_d * h / nu; first-order upwind adds
Pe_d to the diagonal and -Pe_d to the upwind neighbour.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Csr:
    n: int
    ptr: np.ndarray  # int64[n+1]
    col: np.ndarray  # int64[nnz]
    val: np.ndarray  # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.val, self.col, self.ptr), shape=(self.n, self.n))

    def multiply(self, x: np.ndarray) -> np.ndarray:
        return self.to_scipy() @ x


def _from_coo(n: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray) -> Csr:
    """Canonical CSR from duplicate-free COO (rows/cols any order)."""
    keep = vals != 0
    rows, cols, vals = rows[keep], cols[keep], vals[keep]
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ptr, rows + 1, 1)
    np.cumsum(ptr, out=ptr)
    return Csr(n, ptr, cols.astype(np.int64), vals.astype(np.float64))


def _stencil(shape, diag, neighbours) -> Csr:
    """Constant-coefficient stencil on a grid with Dirichlet boundaries.

    shape: grid extents, x fastest.  neighbours: list of (offset vector, value).
    """
    dims = len(shape)
    n = int(np.prod(shape))
    coords = np.indices(shape[::-1]).reshape(dims, -1)[::-1]  # coords[0] = x
    idx = np.arange(n, dtype=np.int64)
    strides = [1]
    for s in shape[:-1]:
        strides.append(strides[-1] * s)
    rows = [idx]
    cols = [idx]
    vals = [np.full(n, float(diag))]
    for off, v in neighbours:
        ok = np.ones(n, dtype=bool)
        shift = 0
        for d in range(dims):
            c = coords[d] + off[d]
            ok &= (c >= 0) & (c < shape[d])
            shift += off[d] * strides[d]
        rows.append(idx[ok])
        cols.append(idx[ok] + shift)
        vals.append(np.full(int(ok.sum()), float(v)))
    return _from_coo(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def convection_diffusion_2d(side: int = 64, pe=(10.0, 5.0)) -> Csr:
    """5-point upwind convection-diffusion, velocity direction (1, 0.5).

    Cell Peclet 10 along x and 5 along y (velocity (1, 0.5) scaled to cell
    Peclet 10); h^2-scaled: diagonal 4+15, west -11, south -6, east/north -1.
    """
    px, py = pe
    return _stencil(
        (side, side),
        4 + px + py,
        [((-1, 0), -1 - px), ((1, 0), -1), ((0, -1), -1 - py), ((0, 1), -1)],
    )


def convection_diffusion_3d(side: int = 128, pe: float = 5.0) -> Csr:
    """7-point upwind convection-diffusion, velocity Pe*(1,1,1), cell Peclet Pe.

    h^2-scaled: diagonal 6+3Pe, upwind neighbours -1-Pe, downwind -1.
    """
    return _stencil(
        (side, side, side),
        6 + 3 * pe,
        [
            ((-1, 0, 0), -1 - pe),
            ((1, 0, 0), -1),
            ((0, -1, 0), -1 - pe),
            ((0, 1, 0), -1),
            ((0, 0, -1), -1 - pe),
            ((0, 0, 1), -1),
        ],
    )


def band(n: int, diagonal: float) -> Csr:
    """Banded family: row i scaled by 2^(i%5-2); a_ii = diagonal, a_i,i-1 = 2,
    a_i,i+1 = 1 before scaling (reference tests/test_cuda_spai_sparse.cpp:7)."""
    i = np.arange(n, dtype=np.int64)
    scale = np.ldexp(1.0, (i % 5) - 2)
    rows = np.concatenate([i, i[1:], i[:-1]])
    cols = np.concatenate([i, i[1:] - 1, i[:-1] + 1])
    vals = np.concatenate([diagonal * scale, 2 * scale[1:], scale[:-1]])
    return _from_coo(n, rows, cols, vals)


def power_law(n: int = 500_000, seed: int = 0x5EED, exponent: float = 2.2, kmin: int = 4,
              kmax: int = 200) -> Csr:
    """Random nonsymmetric matrix with power-law off-diagonal row counts.

    Off-diagonal count per row ~ P(k) proportional to k^-exponent on
    [kmin, kmax]; column indices uniform over the other n-1 columns (repeats
    within a row are merged away, so a row may hold slightly fewer); values
    N(0,1); a_ii = (sum_j |a_ij|) * U(0.6, 1.3), so some rows are not
    diagonally dominant.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    ks = np.arange(kmin, kmax + 1)
    p = ks.astype(np.float64) ** (-exponent)
    p /= p.sum()
    counts = rng.choice(ks, size=n, p=p)
    rows = np.repeat(np.arange(n, dtype=np.int64), counts)
    cols = rng.integers(0, n - 1, size=rows.size, dtype=np.int64)
    cols = cols + (cols >= rows)  # skip the diagonal
    vals = rng.standard_normal(rows.size)
    key = rows * n + cols
    key, first = np.unique(key, return_index=True)
    rows, cols, vals = rows[first], cols[first], vals[first]
    absum = np.zeros(n)
    np.add.at(absum, rows, np.abs(vals))
    diag = absum * rng.uniform(0.6, 1.3, size=n)
    diag[diag == 0] = 1.0
    idx = np.arange(n, dtype=np.int64)
    return _from_coo(n, np.concatenate([rows, idx]), np.concatenate([cols, idx]),
                     np.concatenate([vals, diag]))


def variable_coefficient_2d(side: int, kappa: np.ndarray, pe: float,
                            direction=(1.0, 0.5)) -> Csr:
    """2-D diffusion-convection with cell conductivities kappa (side x side).

    Face conductivity is the arithmetic mean of the two cells; a Dirichlet
    boundary face uses the cell's own value.  Upwind convection with cell
    Peclet pe * direction.
    """
    n = side * side
    k = kappa.reshape(side, side)  # [y, x]
    idx = np.arange(n, dtype=np.int64).reshape(side, side)
    diag = np.zeros((side, side))
    rows, cols, vals = [], [], []
    # x faces
    kf = 0.5 * (k[:, 1:] + k[:, :-1])
    diag[:, 1:] += kf
    diag[:, :-1] += kf
    rows += [idx[:, 1:].ravel(), idx[:, :-1].ravel()]
    cols += [idx[:, :-1].ravel(), idx[:, 1:].ravel()]
    vals += [-kf.ravel(), -kf.ravel()]
    # y faces
    kf = 0.5 * (k[1:, :] + k[:-1, :])
    diag[1:, :] += kf
    diag[:-1, :] += kf
    rows += [idx[1:, :].ravel(), idx[:-1, :].ravel()]
    cols += [idx[:-1, :].ravel(), idx[1:, :].ravel()]
    vals += [-kf.ravel(), -kf.ravel()]
    # boundary faces
    diag[:, 0] += k[:, 0]
    diag[:, -1] += k[:, -1]
    diag[0, :] += k[0, :]
    diag[-1, :] += k[-1, :]
    # upwind convection, velocity components >= 0
    px, py = pe * direction[0], pe * direction[1]
    diag += px + py
    rows += [idx[:, 1:].ravel(), idx[1:, :].ravel()]
    cols += [idx[:, :-1].ravel(), idx[:-1, :].ravel()]
    vals += [np.full(side * (side - 1), -px), np.full(side * (side - 1), -py)]
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    # merge the two x/y off-diagonal contributions of the same (r, c)
    key = r * n + c
    uk, inv = np.unique(key, return_inverse=True)
    acc = np.zeros(uk.size)
    np.add.at(acc, inv, v)
    return _from_coo(n, uk // n, uk % n, acc)


@dataclass
class SystemBatch:
    """Block-diagonal concatenation of independent systems."""

    matrix: Csr
    offsets: np.ndarray  # int64[nsys+1], row range of each system
    rhs: np.ndarray      # float64[sum n], each block equal-component unit norm

    def block(self, s: int) -> Csr:
        """System s as its own CSR (rows and columns renumbered from 0)."""
        a, r0, r1 = self.matrix, int(self.offsets[s]), int(self.offsets[s + 1])
        p0, p1 = int(a.ptr[r0]), int(a.ptr[r1])
        return Csr(r1 - r0, a.ptr[r0:r1 + 1] - p0, a.col[p0:p1] - r0, a.val[p0:p1])


def many_systems(count: int = 256, seed: int = 0xBA7C4, lmin: int = 32, lmax: int = 256,
                 sigma: float = 1.0, pemax: float = 10.0) -> SystemBatch:
    """Independent 2-D variable-coefficient diffusion-convection systems.

    Side L ~ U{lmin..lmax}; cell conductivity log-normal(0, sigma); cell
    Peclet ~ U(0, pemax) along direction (1, 0.5).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = rng.integers(lmin, lmax + 1, size=count)
    blocks = []
    offsets = [0]
    for side in sides:
        kappa = np.exp(sigma * rng.standard_normal((side, side)))
        pe = rng.uniform(0.0, pemax)
        blocks.append(variable_coefficient_2d(int(side), kappa, pe))
        offsets.append(offsets[-1] + int(side) * int(side))
    offsets = np.asarray(offsets, dtype=np.int64)
    return SystemBatch(block_diagonal(blocks), offsets, equal_rhs_blocks(offsets))


def block_diagonal(blocks) -> Csr:
    n = sum(b.n for b in blocks)
    ptr = [np.zeros(1, dtype=np.int64)]
    cols, vals = [], []
    row0 = 0
    nz0 = 0
    for b in blocks:
        ptr.append(b.ptr[1:] + nz0)
        cols.append(b.col + row0)
        vals.append(b.val)
        row0 += b.n
        nz0 += b.nnz
    return Csr(n, np.concatenate(ptr), np.concatenate(cols), np.concatenate(vals))


def equal_rhs(n: int) -> np.ndarray:
    """b_i = 1/sqrt(n), the solver CLI's right side (reference src/solver_main.cpp:94)."""
    return np.full(n, 1.0 / math.sqrt(n))


def equal_rhs_blocks(offsets: np.ndarray) -> np.ndarray:
    out = np.empty(int(offsets[-1]))
    for s in range(len(offsets) - 1):
        a, b = int(offsets[s]), int(offsets[s + 1])
        out[a:b] = 1.0 / math.sqrt(b - a)
    return out


# ------------------------------------------------------------ differentiable A(theta)
@dataclass
class EdgeDiffusion:
    """A(theta) = sum_e theta_e a_e a_e^T on a side x side grid.

    a_e = e_p - e_q for an interior edge and e_p for a Dirichlet boundary edge.
    The CSR values are linear in theta: val = S @ theta, with S stored as
    (value index, edge index, coefficient) triples.
    """

    side: int
    pattern: Csr          # structure of A (values from theta0)
    s_val: np.ndarray     # int64 value-slot index
    s_edge: np.ndarray    # int64 edge index
    s_coef: np.ndarray    # float64 +-1
    n_edges: int

    def values(self, theta: np.ndarray) -> np.ndarray:
        out = np.zeros(self.pattern.nnz)
        np.add.at(out, self.s_val, self.s_coef * theta[self.s_edge])
        return out

    def grad_theta(self, grad_vals: np.ndarray) -> np.ndarray:
        g = np.zeros(self.n_edges)
        np.add.at(g, self.s_edge, self.s_coef * grad_vals[self.s_val])
        return g


def gaussian_field(side_y: int, side_x: int, sigma: float, corr: float,
                   rng: np.random.Generator) -> np.ndarray:
    """Stationary Gaussian field with a Gaussian covariance of length `corr`
    cells, scaled to standard deviation sigma (FFT filtering of white noise)."""
    white = rng.standard_normal((side_y, side_x))
    ky = np.fft.fftfreq(side_y)[:, None]
    kx = np.fft.fftfreq(side_x)[None, :]
    filt = np.exp(-0.5 * (2 * np.pi * corr) ** 2 * (kx ** 2 + ky ** 2))
    f = np.real(np.fft.ifft2(np.fft.fft2(white) * filt))
    f -= f.mean()
    f *= sigma / f.std()
    return f


def diffusion_edges_2d(side: int = 256, seed: int = 0xD1FF, sigma: float = 0.5,
                       corr: float = 8.0):
    """Edge-conductivity diffusion A(theta) with theta = exp(g).

    Returns (EdgeDiffusion, theta, x_star, b): g a Gaussian field (sigma,
    correlation length corr cells) sampled on edges, x* ~ N(0,1), b equal
    components with unit norm.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    n = side * side
    idx = np.arange(n, dtype=np.int64).reshape(side, side)  # [y, x]
    # x faces: face c lies between cells (., c-1) and (., c), c = 0..side;
    # y faces likewise between rows.  A missing cell (-1) is the Dirichlet ghost.
    left = np.full((side, side + 1), -1, dtype=np.int64)
    right = np.full((side, side + 1), -1, dtype=np.int64)
    left[:, 1:] = idx
    right[:, :-1] = idx
    down = np.full((side + 1, side), -1, dtype=np.int64)
    up = np.full((side + 1, side), -1, dtype=np.int64)
    down[1:, :] = idx
    up[:-1, :] = idx
    p = np.concatenate([left.ravel(), down.ravel()])
    q = np.concatenate([right.ravel(), up.ravel()])
    swap = p < 0
    p[swap], q[swap] = q[swap], -1
    n_edges = p.size
    gx = gaussian_field(side, side + 1, sigma, corr, rng)
    gy = gaussian_field(side + 1, side, sigma, corr, rng)
    theta = np.exp(np.concatenate([gx.ravel(), gy.ravel()]))
    # contributions: (p,p,+1), and for interior (q,q,+1), (p,q,-1), (q,p,-1)
    e = np.arange(n_edges, dtype=np.int64)
    inner = q >= 0
    rows = np.concatenate([p, q[inner], p[inner], q[inner]])
    cols = np.concatenate([p, q[inner], q[inner], p[inner]])
    edge = np.concatenate([e, e[inner], e[inner], e[inner]])
    coef = np.concatenate([np.ones(n_edges), np.ones(inner.sum()), -np.ones(inner.sum()),
                           -np.ones(inner.sum())])
    key = rows * n + cols
    uk, slot = np.unique(key, return_inverse=True)
    pattern_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(pattern_ptr, uk // n + 1, 1)
    np.cumsum(pattern_ptr, out=pattern_ptr)
    pattern = Csr(n, pattern_ptr, (uk % n).astype(np.int64), np.zeros(uk.size))
    model = EdgeDiffusion(side, pattern, slot.astype(np.int64), edge, coef, n_edges)
    pattern.val = model.values(theta)
    x_star = rng.standard_normal(n)
    b = equal_rhs(n)
    return model, theta, x_star, b


# ------------------------------------------------------------ Matrix Market
def write_matrix_market(path, a: Csr, comment: str = "") -> None:
    """Coordinate real general, 1-based, values with repr() (round-trip exact)."""
    rows = np.repeat(np.arange(a.n), np.diff(a.ptr))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        if comment:
            f.write(f"% {comment}\n")
        f.write(f"{a.n} {a.n} {a.nnz}\n")
        lines = [f"{r + 1} {c + 1} {v!r}\n" for r, c, v in zip(rows.tolist(), a.col.tolist(),
                                                                 a.val.tolist())]
        f.writelines(lines)
