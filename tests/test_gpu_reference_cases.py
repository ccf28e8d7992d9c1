"""The reference's own edge cases run through the HIP path (C ABI).

  * the 16 RELEASE_CPU_SNAPSHOT cases of the reference
    (tests/release_cpu_snapshot.cpp:285-392, emitted by its compiled sources,
    tests/golden/release_cpu_snapshot.jsonl): SPAI scaling, P, per-column
    patterns, coefficients, fits and residuals bit for bit; GMRES and IR
    statuses, iteration counts and the sequential diagnostics; the nine
    status-only cases;
  * the estimate-versus-recomputed gap (reference tests/test_gmres.cpp:
    222-246): no convergence from the estimate alone;
  * the Matrix Market reject list (reference tests/test_cpu.cpp:67-82 and the
    edge cases of tests/golden/mm_cases.json) through the HIP CLI and through
    read_matrix_market, whose canonicalization runs on the GPU.
Bars: bit-exact where the reference's arithmetic is sequential (SPAI, x0 = P b,
the IR diagnostics of the first iterate); statuses and counts equal; GMRES
solutions within 1e-12 relative (the GPU orthogonalizes in compact form).
"""
import json
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, REPO, bits

pytestmark = pytest.mark.gpu


def _snap():
    with open(os.path.join(GOLDEN, "release_cpu_snapshot.jsonl")) as f:
        return {json.loads(ln.split(" ", 1)[1])["case"]: json.loads(ln.split(" ", 1)[1])
                for ln in f if ln.strip()}


SNAP = _snap()


def fbits(v):
    return np.array([int(s.split("@")[1], 16) for s in v], dtype=np.uint64).view(np.float64)


def csr_of(d):
    from bspai_hip import problems as P

    return P.Csr(d["rows"], np.array(d["offsets"], np.int64), np.array(d["indices"], np.int64),
                 fbits(d["values"]))


def dense(m):
    from bspai_hip import problems as P

    m = np.asarray(m, dtype=float)
    r, c = np.nonzero(m)
    return P._from_coo(m.shape[0], r, c, m[r, c])


@pytest.mark.parametrize("case,eps,alpha,beta", [
    ("spai_direct_full", 1e-12, 2, 8),
    ("spai_tie_growth", 1e-12, 1, 1),
    ("spai_near_dependent", 1e-5, 1, 1),
])
def test_snapshot_construction(ctx, case, eps, alpha, beta):
    import bspai_hip

    g = SNAP[case]
    m = bspai_hip.Matrix(csr_of(g["A"]), ctx)
    s = bspai_hip.construct_spai(m, eps, alpha, beta)
    c = g["construction"]
    assert s.status == c["status"]
    assert np.array_equal(bits(s.scaling()), bits(fbits(c["scaling"])))
    ptr, col, val = s.preconditioner()
    p = c["P"]
    assert list(ptr) == p["offsets"] and list(col) == p["indices"]
    assert np.array_equal(bits(val), bits(fbits(p["values"])))
    pp, pi, pc = s.patterns()
    cols = s.columns()
    for k, colref in enumerate(c["columns"]):
        assert list(pi[pp[k]:pp[k + 1]]) == colref["pattern"]
        assert np.array_equal(bits(pc[pp[k]:pp[k + 1]]), bits(fbits(colref["coefficients"])))
        assert cols["fits"][k] == len(colref["history"])
        if colref["history"]:
            last = colref["history"][-1]
            assert bits([cols["residual"][k]])[0] == bits(fbits([last["residual"]]))[0]


def test_snapshot_gmres_identity(ctx):
    import bspai_hip

    g = SNAP["gmres_identity_nonsymmetric"]
    a = bspai_hip.Matrix(csr_of(g["A"]), ctx)
    p = bspai_hip.Matrix(csr_of(g["fixed_P"]), ctx)
    r = bspai_hip.left_preconditioned_gmres(a, p, fbits(g["b"]), 1e-8, 0)
    ref = g["gmres"]
    assert r.status == ref["status"] and r.iterations == ref["iterations"]
    assert r.converged == ref["converged"]
    xr = fbits(ref["solution"])
    assert np.max(np.abs(np.asarray(r.solution) - xr)) <= 1e-12 * np.max(np.abs(xr))
    # the initial residuals are the sequential ones (same bits)
    assert bits([r.initial_original_residual])[0] == bits(fbits([ref["initial_original"]]))[0]
    assert (bits([r.initial_preconditioned_relative])[0]
            == bits(fbits([ref["initial_preconditioned"]]))[0])
    assert r.final_preconditioned_relative <= 1e-8


@pytest.mark.parametrize("case,kw", [
    ("spai_gmres_ir", dict(epsilon=0.3, alpha=2, beta=8, tolerance=1e-8, max_outer_steps=20)),
    ("ir_identity_preconditioner_correction",
     dict(epsilon=0.3, alpha=2, beta=8, tolerance=1e-8, max_outer_steps=3, fixed=True)),
    ("ir_inner_iteration_limit",
     dict(epsilon=0.3, alpha=2, beta=8, tolerance=1e-8, max_outer_steps=3, max_iterations=1,
          fixed=True)),
])
def test_snapshot_ir(ctx, case, kw):
    import bspai_hip

    g = SNAP[case]
    kw = dict(kw)
    a = bspai_hip.Matrix(csr_of(g["A"]), ctx)
    b = fbits(g["b"])
    if kw.pop("fixed", False):
        p = bspai_hip.Matrix(csr_of(g["fixed_P"]), ctx)
        for k in ("epsilon", "alpha", "beta"):
            kw.pop(k)
        r = bspai_hip.refine_with_preconditioner(a, p, b, **kw)
    else:
        r = bspai_hip.spai_gmres_ir(a, b, **kw)
    ref = g["ir"]
    assert r.status == ref["status"], r.message
    assert r.outer_steps == ref["outer_steps"] and r.inner_attempts == ref["inner_attempts"]
    assert r.inner_statuses == [x["status"] for x in ref["inner_history"]]
    assert r.inner_iterations == [x["iterations"] for x in ref["inner_history"]]
    assert len(r.history) == len(ref["history"])
    # the first iterate x0 = P b and its diagnostics are sequential: same bits
    d0 = ref["history"][0]["diagnostics"]
    exp0 = fbits([d0["inf"], d0["norm2"], d0["nbe"], d0["cbe"]])
    assert bits([r.history[0][0]])[0] == bits(exp0[:1])[0]
    assert bits([r.history[0][2]])[0] == bits(exp0[2:3])[0]
    assert bits([r.history[0][3]])[0] == bits(exp0[3:4])[0]
    assert abs(r.history[0][1] - exp0[1]) <= 1e-14 * exp0[1]
    xr = fbits(ref["solution"])
    assert np.max(np.abs(np.asarray(r.solution) - xr)) <= 1e-12 * max(1.0, np.max(np.abs(xr)))


def test_snapshot_status_only(ctx):
    """The nine status-only snapshot cases (tests/release_cpu_snapshot.cpp:362-392)."""
    import bspai_hip

    M = lambda d: bspai_hip.Matrix(dense(d), ctx)  # noqa: E731
    tiny = M([[4, 1, 0], [2, 5, 1], [0, 1, 3]])
    ident = M(np.eye(3))
    b = np.array([2.0, -5, 7])
    s0 = bspai_hip.construct_spai(tiny, 1e-12, 0, 8)
    r0 = s0.columns()["residual"][0]
    eq = bspai_hip.construct_spai(tiny, r0, 0, 8)
    assert (bspai_hip.api.SPAI_STATUS[eq.columns()["status"][0]]
            == SNAP["spai_epsilon_equal_column0"]["actual_status"])
    below = bspai_hip.construct_spai(tiny, math.nextafter(r0, 0.0), 0, 8)
    assert below.status == SNAP["spai_epsilon_one_ulp_below"]["actual_status"]
    assert (bspai_hip.construct_spai(tiny, 1e-12, 2, 0).status
            == SNAP["spai_beta_zero"]["actual_status"])
    assert (bspai_hip.construct_spai(M([[1, 2], [1, 2]]), 1e-12, 2, 8).status
            == SNAP["spai_rank_zero_pivot"]["actual_status"])
    assert (bspai_hip.construct_spai(M([[1e-320, 0], [0, 1]]), 1e-12, 2, 8).status
            == SNAP["spai_scaling_overflow"]["actual_status"])
    assert (bspai_hip.left_preconditioned_gmres(tiny, ident, np.zeros(3)).status
            == SNAP["gmres_zero_rhs"]["actual_status"])
    assert (bspai_hip.left_preconditioned_gmres(
        tiny, ident, b, initial_guess=np.array([1.0, -2, 3])).status
        == SNAP["gmres_exact_initial"]["actual_status"])
    assert (bspai_hip.left_preconditioned_gmres(tiny, ident, b, 1e-30, 1).status
            == SNAP["gmres_iteration_cap"]["actual_status"])
    assert (bspai_hip.refine_with_preconditioner(tiny, ident, np.zeros(3),
                                                 max_outer_steps=3).status
            == SNAP["ir_zero_rhs"]["actual_status"])


def test_gmres_estimate_gap_never_converges(ctx):
    """The estimate-versus-recomputed gap (reference tests/test_gmres.cpp:
    139-148,222-246) on [[4,1,0],[2,5,1],[0,1,3]], P = I, b = (.1,-.7,1.3).

    At tau = 1e-30 (the reference's strict case) no step can be accepted: the
    full dimension ends in nonconvergence, on the GPU and in the oracle.  At
    tau = 2e-16 the third step's estimate (about 1.1e-16) is below tau and the
    recomputed residual sits within a few ulps of it: the reference's own test
    measured 5.1e-16 on its platform (nonconvergence), the compiled reference
    here gives 1.6e-16 (success), and the GPU's compact Gram-Schmidt lands on
    either side.  What must hold everywhere is the rule the probe exercises:
    the accept decision follows the recomputed residual, so a run whose
    recomputed residual exceeds tau never reports convergence."""
    import bspai_hip
    from oracle import pyoracle as O

    A = dense([[4, 1, 0], [2, 5, 1], [0, 1, 3]])
    I3 = dense(np.eye(3))
    b = np.array([0.1, -0.7, 1.3])
    tiny = bspai_hip.Matrix(A, ctx)
    ident = bspai_hip.Matrix(I3, ctx)
    strict = bspai_hip.left_preconditioned_gmres(tiny, ident, b, 1e-30, 0)
    ref = O.gmres(A, I3, b, 1e-30, 0)
    assert strict.status == ref.status == "nonconvergence" and not strict.converged
    assert strict.iterations == ref.iterations == 3
    for runner in ("gpu", "oracle"):
        r = (bspai_hip.left_preconditioned_gmres(tiny, ident, b, 2e-16, 0) if runner == "gpu"
             else O.gmres(A, I3, b, 2e-16, 0))
        assert r.iterations == 3, runner
        if r.final_preconditioned_relative > 2e-16 if runner == "gpu" else r.final_relative > 2e-16:
            assert r.status == "nonconvergence" and not r.converged, runner
        else:
            assert r.status == "success" and r.converged, runner


MM_CASES = json.load(open(os.path.join(GOLDEN, "mm_cases.json")))
MM_REJECTS = [c for c in MM_CASES if "error" in c["reference"]]


@pytest.mark.parametrize("case", MM_REJECTS, ids=[c["name"] for c in MM_REJECTS])
def test_cli_rejects_like_reference(case):
    """The HIP CLI on each rejected Matrix Market text: the reference CLI's exit
    code and stderr line (runtime errors exit 1, invalid arguments 2)."""
    exe = os.path.join(REPO, "spai-gmres-ir_amd", "hip_fp64_ir")
    with tempfile.TemporaryDirectory() as d:
        mtx = os.path.join(d, "case.mtx")
        with open(mtx, "wb") as f:
            f.write(case["text"].encode())
        r = subprocess.run([exe, mtx], capture_output=True, text=True, timeout=120)
    assert r.returncode == case["cli"]["exit"], r.stderr
    assert r.stderr == case["cli"]["stderr"]
    assert r.stdout == ""


@pytest.mark.parametrize("gpu", [True, False], ids=["gpu-parse", "host-parse"])
@pytest.mark.parametrize("case", MM_CASES, ids=[c["name"] for c in MM_CASES])
def test_read_matrix_market_on_gpu(ctx, case, gpu):
    """read_matrix_market: the text parsed on the GPU (or on the host) and
    canonicalized on the GPU gives the reference reader's CSR bits, or its
    message."""
    import bspai_hip

    ref = case["reference"]
    if "error" in ref:
        kind = bspai_hip.MatrixMarketError if ref["error"] == "runtime_error" else ValueError
        with pytest.raises(kind) as e:
            bspai_hip.read_matrix_market(case["text"].encode(), ctx, gpu=gpu)
        assert str(e.value) == ref["message"]
        return
    rows, cols, ptr, col, val = bspai_hip.read_matrix_market(case["text"].encode(), ctx, gpu=gpu)
    assert (rows, cols) == (ref["rows"], ref["cols"])
    assert list(ptr) == ref["ptr"] and list(col) == ref["col"]
    assert [format(int(x), "016x") for x in np.asarray(val).view(np.uint64)] == ref["val_bits"]
