"""Regenerate the golden fixtures in tests/golden/ from the compiled reference.

Runs in the build container only (it needs the reference tree to build
oracle/_ref via `make -C oracle ref`).  Fixtures written:

  release_cpu_snapshot.jsonl  the reference's 16 RELEASE_CPU_SNAPSHOT cases
                              (tests/release_cpu_snapshot.cpp:281-392), verbatim
  c1_reference.npz            2-D 64x64 convection-diffusion through the
                              reference's spai_gmres_ir (eps 0.3, alpha ceil(n/8),
                              beta 8, tau 1e-8, innercap n, outercap 20): P bits,
                              per-column fits/growths/status, inner iteration
                              counts, diagnostics history bits, solution bits
  band_reference.npz          the reference's banded family n in {16,64,256,1024},
                              diagonal in {8,5} (tests/test_cuda_spai_sparse.cpp:7),
                              eps 0.3, alpha 5, beta 8: P bits and column stats
  cli_tiny.txt                `cpu_fp64_ir tiny_nonsymmetric.mtx 0.5` stdout
                              (tests/fixtures/tiny_nonsymmetric.mtx of the reference)
  cli_c1_head.txt             first 6 lines of `cpu_fp64_ir c1.mtx` stdout + x hash
  cli_dup_overflow.txt        `cpu_fp64_ir dup.mtx` exit code and stderr, duplicate sum > DBL_MAX
  mm_cases.json               Matrix Market texts (the reference's accept and reject
                              lists, tests/test_cpu.cpp:36-82, plus edge cases of its
                              rules) with the reference read_matrix_market's outcome
                              (canonical CSR, or exception kind and message) and, for
                              the rejected ones, the reference CLI's exit code and stderr

`python make_goldens.py mm` writes mm_cases.json only.
"""
from __future__ import annotations

import hashlib
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("BSPAI_REFERENCE", "/root/reference")
sys.path.insert(0, os.path.join(REPO, "spai-gmres-ir_amd"))
from bspai_hip import problems as P  # noqa: E402

BIN = os.path.join(REPO, "oracle", "_ref")


def hex_to_f64(lst):
    return np.array([struct.unpack("<d", bytes.fromhex(h)[::-1])[0] for h in lst],
                    dtype=np.float64)


def run_dump(a, rhs, eps, alpha, beta, tol, maxit, outer, tmp):
    mtx = os.path.join(tmp, "a.mtx")
    out = os.path.join(tmp, "a.json")
    P.write_matrix_market(mtx, a)
    if isinstance(rhs, str):
        rarg = rhs
    else:
        rarg = os.path.join(tmp, "rhs.bin")
        np.ascontiguousarray(rhs, dtype="<f8").tofile(rarg)
    subprocess.run([os.path.join(BIN, "ref_dump"), mtx, rarg, repr(eps), str(alpha), str(beta),
                    repr(tol), str(maxit), str(outer), out], check=True)
    with open(out) as f:
        return json.load(f)


def pack(g, prefix=""):
    cols = g["columns"]
    d = {
        prefix + "spai_status": np.array(g["spai_status"]),
        prefix + "p_ptr": np.array(g["P"]["ptr"], dtype=np.int64),
        prefix + "p_col": np.array(g["P"]["col"], dtype=np.int32),
        prefix + "p_val": hex_to_f64(g["P"]["val"]),
        prefix + "fits": np.array([c["f"] for c in cols], dtype=np.int32),
        prefix + "growths": np.array([c["g"] for c in cols], dtype=np.int32),
        prefix + "clamps": np.array([c["cl"] for c in cols], dtype=np.int32),
        prefix + "col_status": np.array([c["s"] for c in cols]),
        prefix + "residual": hex_to_f64([c["r"] for c in cols]),
        prefix + "scaling": hex_to_f64(g["scaling"]),
    }
    ir = g["ir"]
    d[prefix + "ir_status"] = np.array(ir["status"])
    d[prefix + "ir_outer_steps"] = np.array(ir["outer_steps"])
    d[prefix + "ir_inner_iterations"] = np.array([x["iterations"] for x in ir["inner"]],
                                                 dtype=np.int64)
    d[prefix + "ir_history"] = np.array([hex_to_f64(h) for h in ir["history"]])
    d[prefix + "ir_solution"] = hex_to_f64(ir["solution"])
    return d


DUP_OVERFLOW_MTX = ("%%MatrixMarket matrix coordinate real general\n2 2 3\n"
                    "1 1 1e308\n1 1 1e308\n2 2 1\n")


def cli_dup_overflow(tmp):
    """The reference CLI on a file whose duplicate entries sum past DBL_MAX:
    CsrMatrix::from_entries rejects it (invalid_argument, exit code 2)."""
    mtx = os.path.join(tmp, "dup.mtx")
    with open(mtx, "w") as f:
        f.write(DUP_OVERFLOW_MTX)
    r = subprocess.run([os.path.join(BIN, "cpu_fp64_ir"), mtx], capture_output=True, text=True)
    with open(os.path.join(HERE, "cli_dup_overflow.txt"), "w") as f:
        f.write(f"exit={r.returncode}\n" + r.stdout + r.stderr)


_MM_HEAD = "%%MatrixMarket matrix coordinate real general\n"
MM_CASES = [
    # accepted (reference tests/test_cpu.cpp:36-60)
    ("tiny nonsymmetric", "%%MatrixMarket matrix coordinate real general\n% tiny nonsymmetric "
     "fixture\n3 3 7\n3 3 3\n1 2 1\n2 1 2\n1 1 4\n2 2 5\n3 2 1\n2 3 1\n"),
    ("integer duplicates", "%%MatrixMarket matrix coordinate integer general\n3 4 5\n"
     "1 2 3\n3 4 5\n1 2 -3\n3 4 2\n2 1 0\n"),
    ("symmetric expansion", "%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n"
     "1 1 2\n2 1 3\n2 2 4\n3 2 5\n"),
    ("zero nnz", "%%MatrixMarket matrix coordinate real general\n2 3 0\n"),
    # rejected (reference tests/test_cpu.cpp:67-82)
    ("invalid banner", "not a matrix market file\n"),
    ("array storage", "%%MatrixMarket matrix array real general\n2 2\n1\n"),
    ("complex field", "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 2 3\n"),
    ("pattern field", "%%MatrixMarket matrix coordinate pattern general\n1 1 1\n1 1\n"),
    ("skew symmetry", "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1\n"),
    ("nonsquare symmetric", "%%MatrixMarket matrix coordinate real symmetric\n2 3 0\n"),
    ("zero index", _MM_HEAD + "2 2 1\n0 1 1\n"),
    ("row out of range", _MM_HEAD + "2 2 1\n3 1 1\n"),
    ("column out of range", _MM_HEAD + "2 2 1\n1 3 1\n"),
    ("truncated", _MM_HEAD + "2 2 2\n1 1 1\n"),
    ("trailing", _MM_HEAD + "2 2 1\n1 1 1\n2 2 2\n"),
    ("nan value", _MM_HEAD + "2 2 1\n1 1 NaN\n"),
    ("overflowing value", _MM_HEAD + "2 2 1\n1 1 1e999\n"),
    ("negative count", _MM_HEAD + "2 2 -1\n"),
    # edge cases of the same rules
    ("empty file", ""),
    ("banner only", _MM_HEAD),
    ("short banner", "%%MatrixMarket matrix coordinate real\n1 1 1\n1 1 1\n"),
    ("extra banner field", "%%MatrixMarket matrix coordinate real general x\n1 1 1\n1 1 1\n"),
    ("lowercase magic", "%%matrixmarket matrix coordinate real general\n1 1 1\n1 1 1\n"),
    ("short dimensions", _MM_HEAD + "2 2\n"),
    ("extra dimension field", _MM_HEAD + "2 2 1 7\n1 1 1\n"),
    ("zero rows", _MM_HEAD + "0 2 0\n"),
    ("signed dimension", _MM_HEAD + "+2 2 0\n"),
    ("huge dimension", _MM_HEAD + "99999999999999999999 2 0\n"),
    ("huge index", _MM_HEAD + "2 2 1\n18446744073709551616 1 1\n"),
    ("short entry", _MM_HEAD + "2 2 1\n1 1\n"),
    ("extra entry field", _MM_HEAD + "2 2 1\n1 1 1 2\n"),
    ("word value", _MM_HEAD + "2 2 1\n1 1 abc\n"),
    ("value with suffix", _MM_HEAD + "2 2 1\n1 1 1.5x\n"),
    ("infinite value", _MM_HEAD + "2 2 1\n1 1 inf\n"),
    ("underflowing value", _MM_HEAD + "2 2 1\n1 1 1e-400\n"),
    ("subnormal value", _MM_HEAD + "2 2 1\n1 1 1e-310\n"),
    ("hex value", _MM_HEAD + "2 2 1\n1 2 0x1p3\n"),
    ("integer field exponent", "%%MatrixMarket matrix coordinate integer general\n2 2 1\n"
     "1 1 1e3\n"),
    ("integer field signs", "%%MatrixMarket matrix coordinate integer general\n2 2 2\n"
     "1 1 +7\n2 2 -3\n"),
    ("comments blanks crlf tabs", "%%MatrixMarket matrix coordinate real general\r\n% c\r\n"
     "\r\n  \t\n%% more\n 2\t2 3 \r\n%x\n1 2 0.5\r\n\n\t2 1 -0.25\n   % y\n2 2 4"),
    ("vertical tab line", _MM_HEAD + "\v\n2 2 0\n"),
    ("comment after count", _MM_HEAD + "2 2 1\n1 1 1\n% done\n\n"),
    ("explicit zero and cancellation", _MM_HEAD + "2 2 4\n1 1 0\n2 2 1\n2 2 -1\n1 2 -0.0\n"),
    ("duplicate overflow", _MM_HEAD + "2 2 3\n1 1 1e308\n1 1 1e308\n2 2 1\n"),
    ("symmetric duplicates", "%%MatrixMarket matrix coordinate real symmetric\n2 2 3\n"
     "2 1 1\n1 2 2\n2 2 3\n"),
    ("first error wins", _MM_HEAD + "3 3 3\n1 1 1\n4 1 1\n1 1 x\n"),
]


def run_ref_mm(text, tmp):
    """The reference read_matrix_market on `text`: ("ok", rows, cols, ptr, col, val-bits)
    or ("error", kind, message)."""
    mtx = os.path.join(tmp, "case.mtx")
    out = os.path.join(tmp, "case.bin")
    with open(mtx, "wb") as f:
        f.write(text.encode())
    subprocess.run([os.path.join(BIN, "ref_mm"), mtx, out], check=True)
    raw = open(out, "rb").read()
    kind = struct.unpack("<q", raw[:8])[0]
    if kind:
        return {"error": ["runtime_error", "invalid_argument", "exception"][kind - 1],
                "message": raw[8:].decode()}
    rows, cols, nnz = struct.unpack("<3q", raw[8:32])
    body = np.frombuffer(raw[32:], dtype="<i8")
    ptr = body[: rows + 1]
    col = body[rows + 1: rows + 1 + nnz]
    val = body[rows + 1 + nnz:]
    return {"rows": rows, "cols": cols, "ptr": ptr.tolist(), "col": col.tolist(),
            "val_bits": [format(int(v) & (2**64 - 1), "016x") for v in val]}


def mm_cases(tmp):
    cases = []
    for name, text in MM_CASES:
        c = {"name": name, "text": text, "reference": run_ref_mm(text, tmp)}
        if "error" in c["reference"]:
            mtx = os.path.join(tmp, "cli.mtx")
            with open(mtx, "wb") as f:
                f.write(text.encode())
            r = subprocess.run([os.path.join(BIN, "cpu_fp64_ir"), mtx], capture_output=True,
                               text=True)
            c["cli"] = {"exit": r.returncode, "stderr": r.stderr}
        cases.append(c)
    with open(os.path.join(HERE, "mm_cases.json"), "w") as f:
        json.dump(cases, f, indent=1)
        f.write("\n")


def main():
    if sys.argv[1:] == ["mm"]:
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref",
                        f"REF={REF}"], check=True)
        with tempfile.TemporaryDirectory() as tmp:
            mm_cases(tmp)
        print("mm_cases.json written to", HERE)
        return
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "all", "ref",
                    f"REF={REF}"], check=True)
    snap = subprocess.run([os.path.join(BIN, "cpu_release_snapshot")], check=True,
                          capture_output=True, text=True).stdout
    lines = [ln for ln in snap.splitlines() if ln.startswith("RELEASE_CPU_SNAPSHOT ")]
    assert len(lines) == 16, len(lines)
    with open(os.path.join(HERE, "release_cpu_snapshot.jsonl"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with tempfile.TemporaryDirectory() as tmp:
        a = P.convection_diffusion_2d()
        g = run_dump(a, "equal", 0.3, -(-a.n // 8), 8, 1e-8, 0, 20, tmp)
        np.savez_compressed(os.path.join(HERE, "c1_reference.npz"), **pack(g))
        band = {}
        for diag in (8.0, 5.0):
            for n in (16, 64, 256, 1024):
                b = P.band(n, diag)
                truth = ((np.arange(n) % 7) - 3) / 4.0
                rhs = b.multiply(truth)
                g = run_dump(b, rhs, 0.3, 5, 8, 1e-8, 0, 20, tmp)
                band.update(pack(g, prefix=f"n{n}_d{int(diag)}_"))
                band[f"n{n}_d{int(diag)}_rhs"] = rhs
        np.savez_compressed(os.path.join(HERE, "band_reference.npz"), **band)
        # CLI transcripts
        tiny = os.path.join(REF, "tests", "fixtures", "tiny_nonsymmetric.mtx")
        r = subprocess.run([os.path.join(BIN, "cpu_fp64_ir"), tiny, "0.5"], capture_output=True,
                           text=True)
        with open(os.path.join(HERE, "cli_tiny.txt"), "w") as f:
            f.write(f"exit={r.returncode}\n" + r.stdout)
        mtx = os.path.join(tmp, "c1.mtx")
        P.write_matrix_market(mtx, a)
        r = subprocess.run([os.path.join(BIN, "cpu_fp64_ir"), mtx], capture_output=True,
                           text=True)
        lines = r.stdout.splitlines()
        xs = "\n".join(ln for ln in lines if ln.startswith("x[")) + "\n"
        with open(os.path.join(HERE, "cli_c1_head.txt"), "w") as f:
            f.write(f"exit={r.returncode}\n")
            f.write("\n".join(ln for ln in lines if not ln.startswith("x[")) + "\n")
            f.write("x_sha256=" + hashlib.sha256(xs.encode()).hexdigest() + "\n")
        # an invalid-argument transcript (exit code 2)
        r = subprocess.run([os.path.join(BIN, "cpu_fp64_ir"), mtx, "0.3", "21"],
                           capture_output=True, text=True)
        with open(os.path.join(HERE, "cli_bad_outercap.txt"), "w") as f:
            f.write(f"exit={r.returncode}\n" + r.stdout + r.stderr)
        cli_dup_overflow(tmp)
        mm_cases(tmp)
    print("goldens written to", HERE)


if __name__ == "__main__":
    main()
