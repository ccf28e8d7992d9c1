"""The library's Matrix Market reader (bspai_hip_mm_parse, host code) against
the reference's read_matrix_market (src/cpu.cpp:105-161).

Pins:
  * tests/golden/mm_cases.json: the reference's accept and reject lists
    (tests/test_cpu.cpp:36-82) and edge cases of its rules, with the outcome
    the compiled reference gave (canonical CSR bits, or exception message).
    Accepted texts are canonicalized by the oracle's from_entries (the
    reference's CsrMatrix::from_entries restated), so the parse + canonical
    form is compared end to end.
  * The parse does not depend on how the body is split between threads: the
    same entries (bits) and the same first error for 1..16 threads, with
    errors planted in different byte ranges of a large file.
  * Random texts built from the rules' corner cases, against the compiled
    reference reader (oracle/_ref/ref_mm) when it is present.
No GPU: everything here is host code.
"""
import json
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, REPO

REF_MM = os.path.join(REPO, "oracle", "_ref", "ref_mm")


def _api():
    from bspai_hip import api

    return api


def _canonical(rows, cols, r, c, v):
    """Parse result -> canonical CSR through the oracle's from_entries."""
    from oracle import pyoracle as O

    return O.from_entries(rows, cols, r, c, v)


def _outcome(text, threads=0):
    api = _api()
    try:
        rows, cols, r, c, v = api.parse_matrix_market(text, threads=threads)
    except api.MatrixMarketError as e:
        return {"error": "runtime_error", "message": str(e)}
    try:
        ptr, col, val = _canonical(rows, cols, r, c, v)
    except ValueError as e:
        return {"error": "invalid_argument", "message": str(e)}
    return {"rows": rows, "cols": cols, "ptr": [int(x) for x in ptr],
            "col": [int(x) for x in col],
            "val_bits": [format(int(x), "016x") for x in np.asarray(val).view(np.uint64)]}


CASES = json.load(open(os.path.join(GOLDEN, "mm_cases.json")))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_reference_cases(case):
    got = _outcome(case["text"])
    assert got == case["reference"], case["name"]


def test_reference_reject_list_is_covered():
    """Every rejection of the reference's own test list is among the cases."""
    names = {c["name"] for c in CASES if "error" in c["reference"]}
    for n in ("invalid banner", "array storage", "complex field", "pattern field",
              "skew symmetry", "nonsquare symmetric", "zero index", "row out of range",
              "column out of range", "truncated", "trailing", "nan value",
              "overflowing value", "negative count", "duplicate overflow"):
        assert n in names


def _big_text(n_lines, seed, symmetric=False):
    rng = np.random.default_rng(seed)
    n = 5000
    lines = ["%%MatrixMarket matrix coordinate real " + ("symmetric" if symmetric else "general"),
             "% generated", f"{n} {n} {n_lines}"]
    for i in range(n_lines):
        r = int(rng.integers(1, n + 1))
        c = int(rng.integers(1, r + 1)) if symmetric else int(rng.integers(1, n + 1))
        v = rng.standard_normal() * 10.0 ** int(rng.integers(-5, 6))
        lines.append(f"{r} {c} {v!r}")
        if i % 97 == 0:
            lines.append("% comment inside")
        if i % 131 == 0:
            lines.append("   \t")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("symmetric", [False, True])
def test_thread_split_gives_same_entries(symmetric):
    text = _big_text(60000, 11 + symmetric, symmetric)
    api = _api()
    base = api.parse_matrix_market(text, threads=1)
    for t in (2, 3, 7, 16):
        got = api.parse_matrix_market(text, threads=t)
        assert got[0] == base[0] and got[1] == base[1]
        for a, b in zip(got[2:], base[2:]):
            assert np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))
    # entries in file order: the first data line is the first entry
    first = text.splitlines()[3].split()
    assert base[2][0] == int(first[0]) - 1 and base[3][0] == int(first[1]) - 1
    assert base[4][0] == float(first[2])


def _plant(text, line_no, new):
    lines = text.split("\n")
    lines[line_no] = new
    return "\n".join(lines)


def test_first_error_wins_across_ranges():
    """Two bad lines in different byte ranges: the earlier one's message, for
    every split; a bad line after the last counted entry is trailing data."""
    text = _big_text(40000, 3)
    n_lines = len(text.split("\n"))
    early, late = n_lines // 5, (4 * n_lines) // 5
    t = _plant(_plant(text, early, "1 1 1 1"), late, "0 1 1")
    api = _api()
    for th in (1, 2, 5, 16):
        with pytest.raises(api.MatrixMarketError, match="Extra field in Matrix Market line"):
            api.parse_matrix_market(t, threads=th)
    t2 = _plant(text, late, "1 1 zz")
    for th in (1, 4, 16):
        with pytest.raises(api.MatrixMarketError, match="Invalid coordinate value"):
            api.parse_matrix_market(t2, threads=th)
    # one entry too many: trailing data, even when that extra line is malformed
    t3 = text.rstrip("\n") + "\n1 1 x\n"
    for th in (1, 8):
        with pytest.raises(api.MatrixMarketError, match="Trailing coordinate data"):
            api.parse_matrix_market(t3, threads=th)
    # one entry too few
    body = text.rstrip("\n").split("\n")
    while body[-1].startswith("%") or not body[-1].strip():
        body.pop()
    t4 = "\n".join(body[:-1]) + "\n"
    for th in (1, 8):
        with pytest.raises(api.MatrixMarketError, match="Truncated coordinate data"):
            api.parse_matrix_market(t4, threads=th)


def _ref_outcome(text, tmp):
    mtx = os.path.join(tmp, "f.mtx")
    out = os.path.join(tmp, "f.bin")
    with open(mtx, "wb") as f:
        f.write(text.encode())
    subprocess.run([REF_MM, mtx, out], check=True)
    raw = open(out, "rb").read()
    kind = struct.unpack("<q", raw[:8])[0]
    if kind:
        return {"error": ["runtime_error", "invalid_argument", "exception"][kind - 1],
                "message": raw[8:].decode()}
    rows, cols, nnz = struct.unpack("<3q", raw[8:32])
    body = np.frombuffer(raw[32:], dtype="<i8")
    return {"rows": rows, "cols": cols, "ptr": body[: rows + 1].tolist(),
            "col": body[rows + 1: rows + 1 + nnz].tolist(),
            "val_bits": [format(int(v) & (2**64 - 1), "016x") for v in body[rows + 1 + nnz:]]}


TOKENS_INDEX = ["1", "2", "3", "0", "4", "01", "+1", "-1", "1.0", "99999999999999999999", "x"]
TOKENS_VALUE = ["1", "-2.5", "0", "-0.0", "1e-3", "1e308", "1e309", "1e-320", "nan", "inf",
                "0x1p-2", "1.5e", "+7", "3.", ".5", "1e+2", "abc", "1_0", "-"]


@pytest.mark.skipif(not os.path.exists(REF_MM), reason="compiled reference reader absent")
def test_random_texts_match_reference():
    rng = np.random.default_rng(2024)
    with tempfile.TemporaryDirectory() as tmp:
        for trial in range(300):
            field = rng.choice(["real", "integer", "real", "complex"])
            sym = rng.choice(["general", "symmetric", "general"])
            rows, cols = int(rng.integers(1, 4)), int(rng.integers(1, 4))
            if sym == "symmetric" and rng.random() < 0.8:
                cols = rows
            count = int(rng.integers(0, 5))
            lines = [f"%%MatrixMarket matrix coordinate {field} {sym}"]
            if rng.random() < 0.3:
                lines.append("% comment")
            lines.append(f"{rows} {cols} {count}" + (" 1" if rng.random() < 0.05 else ""))
            for _ in range(count + int(rng.integers(-1, 2))):
                r = rng.choice(TOKENS_INDEX) if rng.random() < 0.2 else str(rng.integers(1, rows + 1))
                c = rng.choice(TOKENS_INDEX) if rng.random() < 0.2 else str(rng.integers(1, cols + 1))
                v = rng.choice(TOKENS_VALUE)
                sep = rng.choice([" ", "\t", "  ", " \v "])
                extra = " 9" if rng.random() < 0.05 else ""
                lines.append(sep.join([r, c, v]) + extra + ("\r" if rng.random() < 0.2 else ""))
                if rng.random() < 0.15:
                    lines.append(rng.choice(["", " ", "%x", "\t% y", "\v"]))
            text = "\n".join(lines) + ("\n" if rng.random() < 0.8 else "")
            want = _ref_outcome(text, tmp)
            for th in (1, 3):
                assert _outcome(text, threads=th) == want, (trial, text)
