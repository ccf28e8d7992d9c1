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


def main():
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
    print("goldens written to", HERE)


if __name__ == "__main__":
    main()
