"""Pins the oracle's canonical CSR from coordinate entries (the reference's
CsrMatrix::from_entries, src/cpu.cpp:48-76) before the GPU path is checked
against it:

  * the reference's own fixtures (tests/test_cpu.cpp:36-60 of the reference,
    transcribed as entries and expected CSR in tests/golden/);
  * brute force on small random inputs: a dictionary that accumulates each
    (row, column) in input order, then sorted keys;
  * the reference itself, compiled from its sources by `make -C oracle ref`
    (oracle/_ref/ref_entries), on random inputs with duplicates, exact
    cancellations, signed zeros and every error: same bits, same message.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, REPO, bits

REF_ENTRIES = os.path.join(REPO, "oracle", "_ref", "ref_entries")


def O():
    from oracle import pyoracle

    return pyoracle


def fixture_cases():
    with open(os.path.join(GOLDEN, "from_entries_cases.json")) as f:
        return json.load(f)


def entry_arrays(entries):
    e = np.array(entries, dtype=object).reshape(-1, 3)
    return (e[:, 0].astype(np.int64), e[:, 1].astype(np.int64),
            np.array([float(v) for v in e[:, 2]], dtype=np.float64))


def random_entries(rng, rows, cols, count, dup=0.4, zero=0.05):
    """Entries with repeated keys (some cancelling exactly) and signed zeros."""
    r = rng.integers(0, rows, count)
    c = rng.integers(0, cols, count)
    v = rng.standard_normal(count) * 10.0 ** rng.integers(-3, 4, count)
    nd = int(dup * count)
    src = rng.integers(0, count, nd)
    dst = rng.integers(0, count, nd)
    r[dst], c[dst] = r[src], c[src]
    canc = dst[: nd // 3]
    v[canc] = -v[src[: nd // 3]]
    nz = rng.integers(0, count, int(zero * count))
    v[nz] = np.where(rng.random(nz.size) < 0.5, 0.0, -0.0)
    return r, c, v


def brute(rows, cols, r, c, v):
    acc = {}
    for i in range(len(v)):
        k = (int(r[i]), int(c[i]))
        acc[k] = acc[k] + v[i] if k in acc else v[i]
    ptr = np.zeros(rows + 1, np.int64)
    col, val = [], []
    for (i, j) in sorted(acc):
        if acc[(i, j)] != 0.0:
            col.append(j)
            val.append(acc[(i, j)])
            ptr[i + 1] += 1
    return np.cumsum(ptr), np.array(col, np.int64), np.array(val, np.float64)


def run_reference(rows, cols, r, c, v):
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([rows, cols, len(v)], np.int64).tofile(f)
            np.asarray(r, np.int64).tofile(f)
            np.asarray(c, np.int64).tofile(f)
            np.asarray(v, np.float64).tofile(f)
        subprocess.run([REF_ENTRIES, fin, fout], check=True)
        raw = open(fout, "rb").read()
    tag = np.frombuffer(raw[:8], np.int64)[0]
    if tag == 1:
        return raw[8:].decode()
    nnz = int(np.frombuffer(raw[8:16], np.int64)[0])
    a = np.frombuffer(raw[16:], np.int64, count=rows + 1 + nnz)
    val = np.frombuffer(raw[16 + 8 * (rows + 1 + nnz):], np.float64, count=nnz)
    return a[: rows + 1], a[rows + 1:], val


def same(a, b):
    return (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            and np.array_equal(bits(a[2]), bits(b[2])))


def test_reference_fixtures():
    d = fixture_cases()
    for case in d["cases"]:
        rows, cols = case["rows"], case["cols"]
        r, c, v = entry_arrays(case["entries"])
        ptr, col, val = O().from_entries(rows, cols, r, c, v)
        assert ptr.tolist() == case["ptr"], case["name"]
        assert col.tolist() == case["col"], case["name"]
        assert val.tolist() == [float(x) for x in case["val"]], case["name"]
        x = np.array(case["multiply"]["x"], np.float64)
        y = np.zeros(rows)
        for i in range(rows):
            for q in range(ptr[i], ptr[i + 1]):
                y[i] += val[q] * x[col[q]]
        assert y.tolist() == [float(t) for t in case["multiply"]["y"]], case["name"]
    for case in d["errors"]:
        r, c, v = entry_arrays(case["entries"])
        with pytest.raises(ValueError, match=case["message"]):
            O().from_entries(case["rows"], case["cols"], r, c, v)


def test_brute_force_small():
    rng = np.random.default_rng(7)
    for trial in range(60):
        rows, cols = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        r, c, v = random_entries(rng, rows, cols, int(rng.integers(0, 40)))
        assert same(O().from_entries(rows, cols, r, c, v), brute(rows, cols, r, c, v)), trial


def test_error_precedence_and_duplicate_overflow():
    # the first offending entry in input order decides; its indices before its value
    with pytest.raises(ValueError, match="must be finite"):
        O().from_entries(2, 2, [0, 5], [0, 0], [np.inf, 1.0])
    with pytest.raises(ValueError, match="outside matrix"):
        O().from_entries(2, 2, [0, 5, 0], [0, 0, 0], [1.0, np.nan, np.nan])
    with pytest.raises(ValueError, match="outside matrix"):
        O().from_entries(2, 2, [7], [0], [np.nan])
    # a finite input whose duplicate sum overflows
    with pytest.raises(ValueError, match="Duplicate sum is nonfinite"):
        O().from_entries(2, 2, [1, 0, 1], [1, 0, 1], [1e308, 1.0, 1e308])
    # index errors are found before any duplicate sum is formed
    with pytest.raises(ValueError, match="outside matrix"):
        O().from_entries(2, 2, [1, 1, 3], [1, 1, 0], [1e308, 1e308, 1.0])


@pytest.mark.skipif(not os.path.exists(REF_ENTRIES),
                    reason="compiled reference not built (make -C oracle ref)")
def test_compiled_reference_random():
    rng = np.random.default_rng(11)
    for trial in range(40):
        rows, cols = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        count = int(rng.integers(0, 4000))
        r, c, v = random_entries(rng, rows, cols, count)
        assert same(O().from_entries(rows, cols, r, c, v), run_reference(rows, cols, r, c, v)), \
            trial
    for r, c, v in (([0, 5], [0, 0], [np.inf, 1.0]), ([0, 5, 0], [0, 0, 0], [1.0, np.nan, 0.0]),
                    ([1, 0, 1], [1, 0, 1], [1e308, 1.0, 1e308]), ([-1], [0], [1.0])):
        with pytest.raises(ValueError) as e:
            O().from_entries(2, 2, r, c, v)
        assert str(e.value) == run_reference(2, 2, np.array(r), np.array(c), np.array(v))
