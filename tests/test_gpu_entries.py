"""Canonical CSR from coordinate entries on the GPU
(bspai_hip_csr_from_entries, the reference's CsrMatrix::from_entries,
src/cpu.cpp:48-76) against the oracle, through the C ABI:

  * the reference's fixtures (tests/test_cpu.cpp:36-60 of the reference);
  * seeded random inputs with duplicates, exact cancellations and signed
    zeros, square and rectangular, host and device arrays: bit-identical
    offsets, columns and values;
  * every error with the reference's message and precedence;
  * the 128^3 convection-diffusion matrix (14.6M entries) shuffled, every
    value split into two halves that sum back exactly: the canonical CSR
    equals the generator's.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, bits

pytestmark = pytest.mark.gpu


def O():
    from oracle import pyoracle

    return pyoracle


def _cases():
    with open(os.path.join(GOLDEN, "from_entries_cases.json")) as f:
        return json.load(f)


def _arrays(entries):
    e = np.array(entries, dtype=object).reshape(-1, 3)
    return (e[:, 0].astype(np.int64), e[:, 1].astype(np.int64),
            np.array([float(v) for v in e[:, 2]], dtype=np.float64))


def _random(rng, rows, cols, count, dup=0.4):
    r = rng.integers(0, rows, count)
    c = rng.integers(0, cols, count)
    v = rng.standard_normal(count) * 10.0 ** rng.integers(-3, 4, count)
    nd = int(dup * count)
    src = rng.integers(0, count, nd)
    dst = rng.integers(0, count, nd)
    r[dst], c[dst] = r[src], c[src]
    v[dst[: nd // 3]] = -v[src[: nd // 3]]
    nz = rng.integers(0, count, count // 20)
    v[nz] = np.where(rng.random(nz.size) < 0.5, 0.0, -0.0)
    return r, c, v


def _same(gpu, ref):
    return (np.array_equal(np.asarray(gpu[0], np.int64), ref[0])
            and np.array_equal(np.asarray(gpu[1], np.int64), ref[1])
            and np.array_equal(bits(gpu[2]), bits(ref[2])))


def test_reference_fixtures(ctx):
    import bspai_hip

    d = _cases()
    for case in d["cases"]:
        r, c, v = _arrays(case["entries"])
        ptr, col, val = bspai_hip.csr_from_entries(case["rows"], case["cols"], r, c, v, ctx)
        assert ptr.tolist() == case["ptr"], case["name"]
        assert col.tolist() == case["col"], case["name"]
        assert val.tolist() == [float(x) for x in case["val"]], case["name"]
    for case in d["errors"]:
        r, c, v = _arrays(case["entries"])
        with pytest.raises(ValueError, match=case["message"]):
            bspai_hip.csr_from_entries(case["rows"], case["cols"], r, c, v, ctx)


@pytest.mark.parametrize("where", ["host", "device"])
def test_random_vs_oracle(ctx, where):
    import torch

    import bspai_hip

    rng = np.random.default_rng(21)
    for trial, (rows, cols, count) in enumerate(
            [(1, 1, 1), (5, 3, 40), (64, 64, 2000), (300, 17, 5000), (17, 300, 5000),
             (1000, 1000, 200000), (4096, 4096, 1 << 20)]):
        r, c, v = _random(rng, rows, cols, count)
        ref = O().from_entries(rows, cols, r, c, v)
        if where == "device":
            dev = torch.device("cuda:0")
            got = bspai_hip.csr_from_entries(rows, cols, torch.from_numpy(r).to(dev),
                                             torch.from_numpy(c).to(dev),
                                             torch.from_numpy(v).to(dev), ctx)
            got = tuple(t.cpu().numpy() for t in got)
        else:
            got = bspai_hip.csr_from_entries(rows, cols, r, c, v, ctx)
        print(f"case {trial}: {rows}x{cols}, {count} entries -> nnz {len(ref[1])}", flush=True)
        assert _same(got, ref), (trial, rows, cols, count)


def test_errors_and_precedence(ctx):
    import bspai_hip

    def msg(rows, cols, r, c, v):
        with pytest.raises(ValueError) as e:
            bspai_hip.csr_from_entries(rows, cols, np.array(r), np.array(c), np.array(v), ctx)
        with pytest.raises(ValueError) as o:
            O().from_entries(rows, cols, r, c, v)
        assert str(e.value) == str(o.value)
        return str(e.value)

    assert msg(2, 2, [0, 5], [0, 0], [np.inf, 1.0]) == "Matrix entries must be finite"
    assert msg(2, 2, [0, 5, 0], [0, 0, 0], [1.0, np.nan, np.nan]) == \
        "Entry index outside matrix dimensions"
    assert msg(2, 2, [7], [0], [np.nan]) == "Entry index outside matrix dimensions"
    assert msg(2, 2, [0], [-1], [1.0]) == "Entry index outside matrix dimensions"
    assert msg(2, 2, [1, 0, 1], [1, 0, 1], [1e308, 1.0, 1e308]) == "Duplicate sum is nonfinite"
    assert msg(2, 2, [1, 1, 3], [1, 1, 0], [1e308, 1e308, 1.0]) == \
        "Entry index outside matrix dimensions"
    # an error deep in a large input: the first offending entry still decides
    rng = np.random.default_rng(3)
    r, c, v = _random(rng, 500, 500, 100000)
    v[70001] = np.nan
    r[90001] = 500
    assert msg(500, 500, r, c, v) == "Matrix entries must be finite"


def test_empty_and_all_cancelling(ctx):
    import bspai_hip

    ptr, col, val = bspai_hip.csr_from_entries(3, 4, np.zeros(0, np.int64),
                                               np.zeros(0, np.int64), np.zeros(0), ctx)
    assert ptr.tolist() == [0, 0, 0, 0] and len(col) == 0 and len(val) == 0
    ptr, col, val = bspai_hip.csr_from_entries(2, 2, [1, 1, 0], [0, 0, 1], [2.5, -2.5, 0.0], ctx)
    assert ptr.tolist() == [0, 0, 0] and len(col) == 0


def test_c2_shuffled_halves(ctx):
    """The 128^3 system's 14,581,760 entries, each split into two halves
    (exact in binary), shuffled: the canonical CSR equals the generator's."""
    import torch

    import bspai_hip
    from bspai_hip import problems as P

    a = P.convection_diffusion_3d()
    n = a.n
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(a.ptr))
    r = np.concatenate([rows, rows])
    c = np.concatenate([a.col, a.col]).astype(np.int64)
    v = np.concatenate([a.val / 2, a.val / 2])
    perm = np.random.default_rng(5).permutation(r.size)
    dev = torch.device("cuda:0")
    ptr, col, val = bspai_hip.csr_from_entries(n, n, torch.from_numpy(r[perm]).to(dev),
                                               torch.from_numpy(c[perm]).to(dev),
                                               torch.from_numpy(v[perm]).to(dev), ctx)
    assert np.array_equal(ptr.cpu().numpy(), np.asarray(a.ptr))
    assert np.array_equal(col.cpu().numpy(), np.asarray(a.col))
    assert np.array_equal(bits(val.cpu().numpy()), bits(a.val))
    m = bspai_hip.Matrix.from_entries(n, torch.from_numpy(r[perm]).to(dev),
                                      torch.from_numpy(c[perm]).to(dev),
                                      torch.from_numpy(v[perm]).to(dev), ctx)
    assert m.n == n and m.nnz == a.nnz
