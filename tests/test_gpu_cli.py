"""The HIP solver CLI against the reference CLI's contract and transcripts
(reference src/solver_main.cpp:60-144, tests/check_solver_cli.py:18-57):
arguments, output lines, exit codes 0/1/2."""
import math
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, REPO

pytestmark = pytest.mark.gpu
BIN = os.path.join(REPO, "spai-gmres-ir_amd")


def run(exe, *args):
    return subprocess.run([os.path.join(BIN, exe), *map(str, args)], capture_output=True,
                          text=True, timeout=300)


def parse(out):
    xs = {int(m.group(1)): float(m.group(2)) for m in re.finditer(r"^x\[(\d+)\]=(\S+)$", out, re.M)}
    kv = dict(re.findall(r"(\w+)=(\S*)", "\n".join(l for l in out.splitlines()
                                                  if not l.startswith("x["))))
    return kv, np.array([xs[i] for i in range(len(xs))])


@pytest.mark.parametrize("exe", ["hip_fp64_ir", "hip_mixed_ir"])
def test_cli_tiny_fixture(exe):
    """[[4,1,0],[2,5,1],[0,1,3]] at eps 0.5: x = [12,2,16]/(50 sqrt 3)
    (reference tests/check_cuda_ir_cli.py:44), exit 0, backward error <= 10u.
    The FP64 flavour takes the reference CLI's outer steps; the mixed flavour's
    FP32 corrections (tau 1e-4) may need one more refinement step."""
    r = run(exe, os.path.join(GOLDEN, "tiny_nonsymmetric.mtx"), "0.5")
    assert r.returncode == 0, r.stderr
    kv, x = parse(r.stdout)
    ref_kv, ref_x = parse(open(os.path.join(GOLDEN, "cli_tiny.txt")).read())
    assert kv["status"] == ref_kv["status"] == "success" and kv["converged"] == "1"
    assert kv["n"] == "3" and kv["nnz"] == "7" and kv["epsilon"] == "0.5"
    if exe == "hip_fp64_ir":
        assert kv["outer_steps"] == ref_kv["outer_steps"] and kv["tau"] == ref_kv["tau"]
    else:
        assert kv["tau"] == "0.0001"
        assert int(ref_kv["outer_steps"]) <= int(kv["outer_steps"]) <= int(ref_kv["outer_steps"]) + 2
    assert float(kv["backward_error"]) <= 10 * 2.0 ** -53
    exact = np.array([12, 2, 16]) / (50 * math.sqrt(3))
    assert np.max(np.abs(x - exact)) <= 1e-15
    assert np.max(np.abs(x - ref_x)) <= 1e-15


def test_cli_c1_matches_reference_transcript():
    from bspai_hip import problems as P

    head = open(os.path.join(GOLDEN, "cli_c1_head.txt")).read()
    ref_kv, _ = parse(head)
    with tempfile.TemporaryDirectory() as d:
        mtx = os.path.join(d, "c1.mtx")
        P.write_matrix_market(mtx, P.convection_diffusion_2d())
        r = run("hip_fp64_ir", mtx)
    assert r.returncode == 0, r.stderr
    kv, x = parse(r.stdout)
    for key in ("n", "nnz", "epsilon", "tau", "status", "converged", "outer_steps"):
        assert kv[key] == ref_kv[key], key
    assert float(kv["backward_error"]) <= 10 * 2.0 ** -53
    assert len(x) == 4096


def test_cli_invalid_arguments_exit_2():
    with tempfile.TemporaryDirectory() as d:
        mtx = os.path.join(d, "t.mtx")
        with open(os.path.join(GOLDEN, "tiny_nonsymmetric.mtx")) as f:
            open(mtx, "w").write(f.read())
        bad = open(os.path.join(GOLDEN, "cli_bad_outercap.txt")).read().splitlines()
        r = run("hip_fp64_ir", mtx, "0.3", "21")
        assert f"exit={r.returncode}" == bad[0] == "exit=2"
        assert r.stderr.strip() == bad[1]
        assert run("hip_fp64_ir").returncode == 2
        assert run("hip_fp64_ir", mtx, "-1").returncode == 2
        assert run("hip_fp64_ir", mtx, "0.3", "5", "4").returncode == 2  # innercap > n
        assert run("hip_fp64_ir", os.path.join(d, "missing.mtx")).returncode == 2


def test_cli_duplicate_overflow_exit_2():
    """Duplicate entries summing past DBL_MAX: the canonicalization (on the
    GPU here) rejects the file with the reference CLI's message and code."""
    bad = open(os.path.join(GOLDEN, "cli_dup_overflow.txt")).read().splitlines()
    with tempfile.TemporaryDirectory() as d:
        mtx = os.path.join(d, "dup.mtx")
        with open(mtx, "w") as f:
            f.write("%%MatrixMarket matrix coordinate real general\n2 2 3\n"
                    "1 1 1e308\n1 1 1e308\n2 2 1\n")
        for exe in ("hip_fp64_ir", "hip_mixed_ir"):
            r = run(exe, mtx)
            assert f"exit={r.returncode}" == bad[0] == "exit=2"
            assert r.stderr.strip() == bad[1]


def test_cli_failure_exit_1():
    """A construction failure (zero diagonal) is a solver failure: exit 1."""
    with tempfile.TemporaryDirectory() as d:
        mtx = os.path.join(d, "z.mtx")
        with open(mtx, "w") as f:
            f.write("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 2 1\n2 1 1\n2 2 1\n")
        r = run("hip_fp64_ir", mtx)
    assert r.returncode == 1
    kv, _ = parse(r.stdout)
    assert kv["status"] == "construction_failure" and kv["converged"] == "0"


REF_CLI_HIP = os.path.join(REPO, "oracle", "_ref", "ref_cli_hip")


@pytest.mark.skipif(not os.path.exists(REF_CLI_HIP),
                    reason="reference CLI with the HIP branch not built (needs the reference tree)")
def test_reference_cli_with_hip_backend():
    """The reference's own solver CLI (src/solver_main.cpp) with the
    BSPAI_SOLVER_HIP branch a maintainer adds, linked against the reference's
    CPU sources, this repository's C++ adapter and libbspai_hip.so: the drop-in
    in use.  Its transcripts match the reference CPU CLI's (tiny fixture: the
    exact solution; the 64x64 system: the header lines and outer steps)."""
    from bspai_hip import problems as P

    r = subprocess.run([REF_CLI_HIP, os.path.join(GOLDEN, "tiny_nonsymmetric.mtx"), "0.5"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    kv, x = parse(r.stdout)
    ref_kv, ref_x = parse(open(os.path.join(GOLDEN, "cli_tiny.txt")).read())
    for key in ("n", "nnz", "epsilon", "tau", "status", "converged", "outer_steps"):
        assert kv[key] == ref_kv[key], key
    assert kv["backend"] == "GPU" and kv["precision"] == "FP64"
    assert np.max(np.abs(x - ref_x)) <= 1e-15
    head = open(os.path.join(GOLDEN, "cli_c1_head.txt")).read()
    ref_kv, _ = parse(head)
    with tempfile.TemporaryDirectory() as d:
        mtx = os.path.join(d, "c1.mtx")
        P.write_matrix_market(mtx, P.convection_diffusion_2d())
        r = subprocess.run([REF_CLI_HIP, mtx], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    kv, x = parse(r.stdout)
    for key in ("n", "nnz", "epsilon", "tau", "status", "converged", "outer_steps"):
        assert kv[key] == ref_kv[key], key
    assert float(kv["backward_error"]) <= 10 * 2.0 ** -53
    assert len(x) == 4096
