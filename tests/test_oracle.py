"""Pins the CPU oracle (oracle/bspai_oracle.cpp) before anything is checked
against it.

  * the reference's 16 release-snapshot cases (tests/release_cpu_snapshot.cpp
    of the reference, emitted by its compiled sources): exact bits of SPAI
    scaling, P, per-column patterns/coefficients, per-fit shadows, candidates,
    scores and selections; GMRES solutions and residuals; IR iterates and
    diagnostics; status-only edge cases;
  * the 64x64 convection-diffusion system run through the reference's
    cpu_fp64_ir path (tests/golden/c1_reference.npz): P bits, column work, inner
    iteration counts, diagnostics history and solution bits;
  * the reference's banded SPAI family (tests/golden/band_reference.npz);
  * closed forms: the 3x3 rational inverse, diagonal inverses, the sqrt(3.5)
    score and its 32u clamp boundary (reference tests/test_spai.cpp:139-250).
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, bits


def P():
    from bspai_hip import problems

    return problems


def O():
    from oracle import pyoracle

    return pyoracle


def snapshot_cases():
    with open(os.path.join(GOLDEN, "release_cpu_snapshot.jsonl")) as f:
        return {json.loads(ln.split(" ", 1)[1])["case"]: json.loads(ln.split(" ", 1)[1])
                for ln in f if ln.strip()}


def fbits(v):
    """'0x1.8p1@4008000000000000' -> float via the raw IEEE bits."""
    return np.array([int(s.split("@")[1], 16) for s in v], dtype=np.uint64).view(np.float64)


def csr_of(d):
    return P().Csr(d["rows"], np.array(d["offsets"], np.int64), np.array(d["indices"], np.int64),
                   fbits(d["values"]))


SNAP = snapshot_cases()


def test_snapshot_has_16_cases():
    assert len(SNAP) == 16


def _check_construction(case, eps, alpha, beta):
    g = SNAP[case]
    a = csr_of(g["A"])
    res = O().construct_spai(a, eps, alpha, beta, trace=True)
    c = g["construction"]
    assert res.status == c["status"]
    assert np.array_equal(bits(res.scaling), bits(fbits(c["scaling"])))
    p = c["P"]
    assert np.array_equal(res.p_ptr, p["offsets"])
    assert np.array_equal(res.p_col, p["indices"])
    assert np.array_equal(bits(res.p_val), bits(fbits(p["values"])))
    for k, col in enumerate(c["columns"]):
        lo, hi = res.pat_ptr[k], res.pat_ptr[k + 1]
        assert list(res.pat_idx[lo:hi]) == col["pattern"]
        assert np.array_equal(bits(res.pat_coef[lo:hi]), bits(fbits(col["coefficients"])))
        assert res.fits[k] == len(col["history"])
        for f, fit in enumerate(col["history"]):
            assert list(res.trace(k, f, "shadow")) == fit["I"]
            if fit["candidates"]:
                assert list(res.trace(k, f, "candidates")) == fit["candidates"]
                assert np.array_equal(bits(res.trace_scores(k, f)), bits(fbits(fit["scores"])))
                assert list(res.trace(k, f, "selected")) == fit["selected"]
        if col["history"]:
            last = col["history"][-1]
            assert bits([res.residual[k]])[0] == bits(fbits([last["residual"]]))[0]
            # the full fitted residual is the reduced one bit for bit
            assert last["residual"] == last["full_residual"]


def test_snapshot_spai_direct_full():
    _check_construction("spai_direct_full", 1e-12, 2, 8)


def test_snapshot_spai_tie_growth():
    _check_construction("spai_tie_growth", 1e-12, 1, 1)


def test_snapshot_spai_near_dependent():
    _check_construction("spai_near_dependent", 1e-5, 1, 1)


def test_snapshot_gmres_identity():
    g = SNAP["gmres_identity_nonsymmetric"]
    a = csr_of(g["A"])
    p = csr_of(g["fixed_P"])
    r = O().gmres(a, p, fbits(g["b"]), 1e-8, 0)
    ref = g["gmres"]
    assert r.status == ref["status"] and r.iterations == ref["iterations"]
    assert r.converged == ref["converged"]
    assert np.array_equal(bits(r.solution), bits(fbits(ref["solution"])))
    assert bits([r.final_original])[0] == bits(fbits([ref["final_original"]]))[0]
    assert bits([r.final_relative])[0] == bits(fbits([ref["final_preconditioned"]]))[0]
    assert bits([r.initial_relative])[0] == bits(fbits([ref["initial_preconditioned"]]))[0]


@pytest.mark.parametrize("case,kw", [
    ("spai_gmres_ir", dict(epsilon=0.3, alpha=2, beta=8, tolerance=1e-8, max_outer_steps=20)),
    ("ir_identity_preconditioner_correction",
     dict(epsilon=0.3, alpha=2, beta=8, tolerance=1e-8, max_outer_steps=3, fixed=True)),
    ("ir_inner_iteration_limit",
     dict(epsilon=0.3, alpha=2, beta=8, tolerance=1e-8, max_outer_steps=3, max_iterations=1,
          fixed=True)),
])
def test_snapshot_ir(case, kw):
    g = SNAP[case]
    a = csr_of(g["A"])
    kw = dict(kw)
    p = csr_of(g["fixed_P"]) if kw.pop("fixed", False) else None
    r = O().spai_gmres_ir(a, fbits(g["b"]), p=p, **kw)
    ref = g["ir"]
    assert r.status == ref["status"]
    assert r.outer_steps == ref["outer_steps"] and r.inner_attempts == ref["inner_attempts"]
    assert np.array_equal(bits(r.solution), bits(fbits(ref["solution"])))
    assert len(r.history) == len(ref["history"])
    for h, hr in zip(r.history, ref["history"]):
        d = hr["diagnostics"]
        exp = fbits([d["inf"], d["norm2"], d["nbe"], d["cbe"]])
        assert np.array_equal(bits(h), bits(exp))
    assert [int(x) for x in r.inner[:, 2]] == [x["iterations"] for x in ref["inner_history"]]


def _dense(m):
    m = np.asarray(m, dtype=float)
    r, c = np.nonzero(m)
    return P()._from_coo(m.shape[0], r, c, m[r, c])


def test_snapshot_status_only():
    tiny = _dense([[4, 1, 0], [2, 5, 1], [0, 1, 3]])
    ident = _dense(np.eye(3))
    b = np.array([2.0, -5, 7])
    s0 = O().construct_spai(tiny, 1e-12, 0, 8)
    eq = O().construct_spai(tiny, s0.residual[0], 0, 8)
    assert O().SPAI_STATUS[eq.col_status[0]] == SNAP["spai_epsilon_equal_column0"]["actual_status"]
    below = O().construct_spai(tiny, math.nextafter(s0.residual[0], 0.0), 0, 8)
    assert below.status == SNAP["spai_epsilon_one_ulp_below"]["actual_status"]
    assert O().construct_spai(tiny, 1e-12, 2, 0).status == SNAP["spai_beta_zero"]["actual_status"]
    assert (O().construct_spai(_dense([[1, 2], [1, 2]]), 1e-12, 2, 8).status
            == SNAP["spai_rank_zero_pivot"]["actual_status"])
    assert (O().construct_spai(_dense([[1e-320, 0], [0, 1]]), 1e-12, 2, 8).status
            == SNAP["spai_scaling_overflow"]["actual_status"])
    assert O().gmres(tiny, ident, np.zeros(3)).status == SNAP["gmres_zero_rhs"]["actual_status"]
    assert (O().gmres(tiny, ident, b, x0=np.array([1.0, -2, 3])).status
            == SNAP["gmres_exact_initial"]["actual_status"])
    assert (O().gmres(tiny, ident, b, 1e-30, 1).status
            == SNAP["gmres_iteration_cap"]["actual_status"])
    assert (O().spai_gmres_ir(tiny, np.zeros(3), p=ident, max_outer_steps=3).status
            == SNAP["ir_zero_rhs"]["actual_status"])


def test_c1_reference_bit_exact(c1_golden):
    a = P().convection_diffusion_2d()
    b = P().equal_rhs(a.n)
    s = O().construct_spai(a, 0.3)
    assert s.status == str(c1_golden["spai_status"])
    assert np.array_equal(s.fits, c1_golden["fits"])
    assert np.array_equal(s.growths, c1_golden["growths"])
    assert np.array_equal(bits(s.residual), bits(c1_golden["residual"]))
    assert np.array_equal(s.p_ptr, c1_golden["p_ptr"])
    assert np.array_equal(s.p_col, c1_golden["p_col"])
    assert np.array_equal(bits(s.p_val), bits(c1_golden["p_val"]))
    r = O().spai_gmres_ir(a, b, 0.3, p=P().Csr(a.n, s.p_ptr, s.p_col, s.p_val))
    assert r.status == str(c1_golden["ir_status"])
    assert r.outer_steps == int(c1_golden["ir_outer_steps"])
    assert r.inner_iterations == list(c1_golden["ir_inner_iterations"])
    assert np.array_equal(bits(r.history), bits(c1_golden["ir_history"]))
    assert np.array_equal(bits(r.solution), bits(c1_golden["ir_solution"]))


@pytest.mark.parametrize("n", [16, 64, 256, 1024])
@pytest.mark.parametrize("diag", [8, 5])
def test_band_reference_bit_exact(band_golden, n, diag):
    pre = f"n{n}_d{diag}_"
    g = {k[len(pre):]: v for k, v in band_golden.items() if k.startswith(pre)}
    a = P().band(n, float(diag))
    r = O().spai_gmres_ir(a, g["rhs"], 0.3, alpha=5, beta=8)
    assert r.spai_status == str(g["spai_status"])
    pr, pc, pv = r.p
    assert np.array_equal(pr, g["p_ptr"]) and np.array_equal(pc, g["p_col"])
    assert np.array_equal(bits(pv), bits(g["p_val"]))
    assert r.status == str(g["ir_status"])
    assert r.inner_iterations == list(g["ir_inner_iterations"])
    assert np.array_equal(bits(r.solution), bits(g["ir_solution"]))


def test_closed_form_inverse_and_diagonal():
    """3x3 exact rational inverse /50 and diagonal inverses (reference tests/test_spai.cpp:145,203-207)."""
    s = O().construct_spai(_dense([[4, 1, 0], [2, 5, 1], [0, 1, 3]]), 1e-12, 2, 8)
    dense = np.zeros((3, 3))
    for r in range(3):
        dense[r, s.p_col[s.p_ptr[r]:s.p_ptr[r + 1]]] = s.p_val[s.p_ptr[r]:s.p_ptr[r + 1]]
    exact = np.array([[14, -3, 1], [-6, 12, -4], [2, -4, 18]]) / 50.0
    assert np.max(np.abs(dense - exact)) <= 768 * 2.0 ** -52
    d = O().construct_spai(_dense([[2, 0], [0, -4]]), 1e-12, 0, 8)
    assert d.status == "success" and list(d.p_val) == [0.5, -0.25]


def test_score_known_answers():
    st, sc, cl = O().score(4, 1, 2)
    assert st == "success" and sc == math.sqrt(3.5) and not cl
    st, sc, cl = O().score(1, math.nextafter(1.0, 2.0), 1)
    assert st == "success" and cl and sc == 0
    bd = 1 + 16 * 2.0 ** -52
    assert O().score(1, bd, 1)[0] == "success"
    assert O().score(1, math.nextafter(bd, 2.0), 1)[0] == "invalid_score"
    assert O().score(1, 1, 0)[0] == "zero_denominator"
    assert O().score(-1, 1, 1)[0] == "invalid_score"


def test_subset_matches_full_build():
    a = P().convection_diffusion_2d()
    s = O().construct_spai(a, 0.3)
    cols = np.array([0, 1, 63, 64, 65, 2047, 4000, 4095])
    st, fi, gr, pt, ix, cf = O().spai_subset(a, cols)
    for i, k in enumerate(cols):
        lo, hi = s.pat_ptr[k], s.pat_ptr[k + 1]
        assert np.array_equal(ix[pt[i]:pt[i + 1]], s.pat_idx[lo:hi])
        assert np.array_equal(bits(cf[pt[i]:pt[i + 1]]), bits(s.pat_coef[lo:hi]))
        assert fi[i] == s.fits[k] and gr[i] == s.growths[k]


def test_gmres_brute_force_small():
    """Against a dense direct solve: P A x = P b on a 30x30 random system."""
    rng = np.random.default_rng(11)
    m = rng.standard_normal((30, 30)) + 10 * np.eye(30)
    a = _dense(m)
    s = O().construct_spai(a, 0.3)
    p = P().Csr(30, s.p_ptr, s.p_col, s.p_val)
    b = rng.standard_normal(30)
    r = O().gmres(a, p, b, 1e-12, 0)
    assert r.status == "success"
    assert np.allclose(r.solution, np.linalg.solve(m, b), rtol=1e-9, atol=1e-12)
