"""SPAI-GMRES-IR benchmark on one MI355X.

One step = one complete solve of the 3-D 128^3 convection-diffusion system
(n = 2,097,152, nnz = 14,581,760; 7-point upwind, cell Peclet 5): SPAI
construction (epsilon 0.3, alpha ceil(n/8), beta 8) followed by iterative
refinement until ||b-Ax||_2/||b||_2 <= 1e-12 and the reference's backward-error
stop (outercap 20, innercap 200).  `value` is the mixed FP32/FP64 path's
solves/s; the plain FP64 path is timed in the same run and reported beside it.
Inputs (matrix, right side) are resident in HBM before the timed region.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--extra many,powerlaw,canon]
       python bench.py --pmc-workload      (the short run profiled for HBM traffic)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "spai-gmres-ir_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-c1-trials", type=int, default=3,
                    help="warm-up and measured trials of the reference on the 64x64 system")
    ap.add_argument("--extra", default="",
                    help="comma list of extra workloads to time beside the headline: "
                         "'many' (256 systems batched vs looped), 'powerlaw' (uneven columns), "
                         "'canon' (canonical CSR from 29M coordinate entries)")
    ap.add_argument("--powerlaw-n", type=int, default=50000,
                    help="rows of the power-law matrix timed by --extra powerlaw")
    ap.add_argument("--only-extra", action="store_true",
                    help="time only the --extra workloads (no headline, no CPU baseline) and "
                         "print their JSON line")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="run the short workload profiled for HBM traffic counters instead of "
                         "the benchmark: a 1 GiB device copy (counter calibration), one "
                         "SPAI build and 150 GMRES iterations on the same system")
    ap.add_argument("--pmc-precision", default="mixed", choices=["mixed", "fp64"],
                    help="precision of the --pmc-workload build and iterations")
    return ap.parse_args()


# HBM traffic per launch of a kernel on this workload, measured by rocprofv3
# FETCH_SIZE / WRITE_SIZE passes over `bench.py --pmc-workload` and scaled by
# the counters' reading for a copy of known size (the profile states both).
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "spai_gmres_kernels_c2_128cube_hbm_traffic.md")


def measured_traffic(kernel):
    import re

    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    for name, val in re.findall(r"^traffic_bytes_per_launch (.+) (\S+)$", text, re.M):
        if name == kernel:
            return float(val)
    return None


def pmc_workload(args):
    import numpy as np
    import torch

    import bspai_hip
    from bspai_hip import problems as P

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(dev)
    torch.cuda.set_stream(stream)
    ctx = bspai_hip.Context(0, stream=stream.cuda_stream)
    src = torch.ones(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    a = P.convection_diffusion_3d(args.side)
    m = bspai_hip.Matrix(a, ctx)
    prec = args.pmc_precision
    s = bspai_hip.construct_spai(m, 0.3, -(-a.n // 8), 8, precision=prec)
    b = torch.from_numpy(P.equal_rhs(a.n)).to(dev)
    # 150 Krylov steps (k = 0..149, mean 74.5): one outer step of the
    # benchmark's solve, so that per-launch counts compare with its averages.
    r = bspai_hip.refine_with_preconditioner(m, s, b, max_iterations=150, max_outer_steps=1,
                                             precision=prec, tolerance=1e-30)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_workload": "done", "precision": prec, "spai": s.status,
                      "inner_iterations": r.inner_iterations}))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path "
              "is a design note, not a runnable configuration.", file=sys.stderr)
        sys.exit(2)
    if args.pmc_workload:
        pmc_workload(args)
        return
    import numpy as np
    import torch

    import bspai_hip
    from bspai_hip import problems as P

    dev = torch.device("cuda:0")
    # A dedicated stream (torch's default stream has handle 0, which the
    # library would replace by a stream of its own): the library's kernels and
    # the timing events below are then on the same stream.
    stream = torch.cuda.Stream(dev)
    torch.cuda.set_stream(stream)
    ctx = bspai_hip.Context(0, stream=stream.cuda_stream)
    if args.only_extra:
        extras = [e for e in args.extra.split(",") if e]
        out = {}
        if "many" in extras:
            out["many_systems"] = bench_many(ctx)
        if "powerlaw" in extras:
            out["power_law"] = bench_power_law(ctx, args.powerlaw_n)
        if "canon" in extras:
            out["canonical_csr"] = bench_canon(ctx, cpu=False)
        print(json.dumps({"extra": out}), flush=True)
        return
    a = P.convection_diffusion_3d(side=args.side)
    n, nnz = a.n, a.nnz
    b_host = P.equal_rhs(n)
    m = bspai_hip.Matrix(a, ctx)
    b = torch.from_numpy(b_host).to(dev)
    A_sp = a.to_scipy()
    common = dict(epsilon=0.3, alpha=0, beta=8, max_iterations=200, max_outer_steps=20,
                  relres_target=1e-12)
    # Inner tolerances.  The FP32 correction GMRES cannot go below about
    # kappa(PA) u32 (relative, on ||P(r - Ad)||/||Pr||): on this system a
    # tolerance of 5e-6 or less never verifies and the first inner solve
    # reaches the cap of 200.  The mixed path's tolerance is therefore a
    # floor, 2^-16 = 256 FP32 units (1.53e-5), and the library loosens it for
    # a refinement step that needs less contraction to meet the outer stop
    # (adaptive rule, include/bspai_hip.h).  Three refinement steps reach the
    # backward-error stop (eta <= 10u = 1.1e-15) with a 3x margin (eta
    # 3.4e-16); measured: 2e-5 gives a 2.3x margin, 3e-5 only 1.35x.
    # FP64 uses the reference's default 1e-8 (include/bspai/gmres.hpp:24).
    paths = {"mixed": dict(precision="mixed", tolerance=2.0 ** -16),
             "fp64": dict(precision="fp64", tolerance=1e-8)}
    results, kernels, live, solutions = {}, {}, {}, {}
    for name, kw in paths.items():
        last = None
        for _ in range(args.warmup):
            last = bspai_hip.spai_gmres_ir(m, b, **common, **kw)
        torch.cuda.synchronize()
        # HIP-event timing of the SPAI construction kernels over the timed
        # region (events on the library's stream, which is this stream; two
        # events per launch cost nothing measurable).
        ctx.set_timing("spai")
        t0 = time.perf_counter()
        runs = []
        for _ in range(args.steps):
            runs.append(bspai_hip.spai_gmres_ir(m, b, **common, **kw))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        live[name] = ctx.kernel_times()
        # Every hot kernel timed, on one more solve after the timed region
        # (bracketing each of the ~10 launches per Arnoldi step with events
        # costs a few percent, which stays out of the headline).
        ctx.set_timing(True)
        bspai_hip.spai_gmres_ir(m, b, **common, **kw)
        kernels[name] = ctx.kernel_times()
        ctx.set_timing(False)
        last = runs[-1]
        x = last.solution.cpu().numpy()
        solutions[name] = x
        relres = float(np.linalg.norm(b_host - A_sp @ x) / np.linalg.norm(b_host))
        ok = all(r.status == "success" for r in runs) and relres <= 1e-12
        results[name] = dict(
            solves_per_s=args.steps / el, ms_per_solve=1e3 * el / args.steps, status=last.status,
            converged_all=ok, relres=relres, nbe=last.normwise_backward_error,
            outer_steps=last.outer_steps, inner_iterations=last.inner_iterations,
            spai_ms=last.spai_ms, solve_ms=last.solve_ms,
            gmres_iterations_per_s=sum(last.inner_iterations) / (el / args.steps))
    # ---- rooflines: the per-kernel table of the headline (mixed) path names
    # the dominant kernel; when that is the SPAI construction, its figures
    # are the ones timed live over the timed region.
    roof = {name: kernel_rooflines(kt) for name, kt in kernels.items()}
    dominant = max(roof["mixed"], key=lambda r: r["total_ms"])
    live_rows = {r["kernel"]: r for r in kernel_rooflines(live["mixed"])}
    timed_live = dominant["kernel"] in live_rows
    if timed_live:
        dominant = live_rows[dominant["kernel"]]
    roofline = {k: dominant[k] for k in ("kernel", "bound", "achieved", "peak", "unit", "frac")}
    traffic = measured_traffic(dominant["kernel"])
    roofline.update(traffic=traffic, per_launch=dominant["per_launch"], avg_launch_us=dominant["avg_us"],
                    launches=dominant["launches"],
                    measured="HIP events over the timed region" if timed_live else
                    "HIP events, one extra solve after the timed region")
    if traffic is not None:
        roofline["traffic_unit"] = "bytes per launch (HBM, rocprofv3 FETCH_SIZE/WRITE_SIZE)"
        roofline["traffic_source"] = os.path.relpath(TRAFFIC_PROFILE, REPO)
    if timed_live:
        roofline["share_of_timed_region"] = round(
            dominant["total_ms"] / (results["mixed"]["ms_per_solve"] * args.steps), 4)
    # ---- the FP32 preconditioner against the FP64 one (same device build
    # the benchmark solves with)
    s64 = bspai_hip.construct_spai(m, 0.3, -(-n // 8), 8)
    s32 = bspai_hip.construct_spai(m, 0.3, -(-n // 8), 8, precision="mixed")
    p64, p32 = s64.preconditioner(), s32.preconditioner()
    ndiff, colerr = compare_rows(p32, p64)
    precision_report = {"fp32_spai_vs_fp64_spai": {
        "pattern_differing_columns": ndiff, "columns": n,
        "max_column_rel_error_on_common_patterns": colerr},
        "inner_iterations": {"fp64": results["fp64"]["inner_iterations"],
                             "mixed": results["mixed"]["inner_iterations"]}}
    # ---- CPU baseline: one full oracle solve of the same system, timed
    cpu = cpu_c1 = None
    if not args.no_cpu:
        cpu, oracle_check = cpu_baseline(a, b_host, p64, p32, solutions)
        precision_report["against_cpu_oracle"] = oracle_check
        cpu_c1 = cpu_baseline_c1(args.cpu_c1_trials)
    mixed, fp64 = results["mixed"], results["fp64"]
    line = {
        "metric": "solves/s to rel-res 1e-12: mixed FP32/FP64 vs plain FP64 GPU; SpMV % HBM peak",
        "value": round(mixed["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_solve"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (deterministic stencil generator)",
        "config": {"workload": f"3-D convection-diffusion {args.side}^3, n={n}, nnz={nnz}, "
                               "eps=0.3, outercap=20, innercap=200, rel-res 1e-12",
                   "parallelism": "single GPU"},
        "plain_fp64": {k: fp64[k] for k in ("solves_per_s", "ms_per_solve", "status", "relres",
                                            "nbe", "outer_steps", "inner_iterations",
                                            "spai_ms", "solve_ms", "gmres_iterations_per_s")},
        "mixed": {k: mixed[k] for k in ("solves_per_s", "ms_per_solve", "status", "relres",
                                        "nbe", "outer_steps", "inner_iterations", "spai_ms",
                                        "solve_ms", "gmres_iterations_per_s")},
        "mixed_over_fp64": round(mixed["solves_per_s"] / fp64["solves_per_s"], 3),
        "all_converged": mixed["converged_all"] and fp64["converged_all"],
        "roofline": roofline,
        "kernels": {name: [{k: r[k] for k in ("kernel", "launches", "avg_us", "share", "bound",
                                              "achieved", "unit", "frac")}
                           for r in rows[:8]] for name, rows in roof.items()},
        # the SpMV rows on their own (>= 70% of HBM peak),
        # against the algorithmic bytes of an int32-column CSR product
        # (value + 4-byte column per entry, row offsets, x and y once)
        "spmv": {name: {r["kernel"]: {"avg_us": r["avg_us"], "achieved_GBs": r["achieved"],
                                      "frac": r["frac"]}
                        for r in rows if r["kernel"].startswith("spmv_")}
                 for name, rows in roof.items()},
        "cpu_baseline": cpu,
        "cpu_baseline_c1": cpu_c1,
        "precision_report": precision_report,
    }
    extras = [e for e in args.extra.split(",") if e]
    if extras:
        line["extra"] = {}
    if "many" in extras:
        line["extra"]["many_systems"] = bench_many(ctx)
    if "powerlaw" in extras:
        line["extra"]["power_law"] = bench_power_law(ctx, args.powerlaw_n)
    if "canon" in extras:
        line["extra"]["canonical_csr"] = bench_canon(ctx, cpu=not args.no_cpu)
    print(json.dumps(line))


def bench_many(ctx):
    """256 independent variable-coefficient systems (L in 32..256): one batched
    SPAI + IR solve of all of them against the single-system plain path looped
    over the systems; aggregate solves/s."""
    import numpy as np

    import bspai_hip
    from bspai_hip import problems as P

    sb = P.many_systems()
    ns = len(sb.offsets) - 1
    mb = bspai_hip.Matrix(sb.matrix, ctx, systems=sb.offsets)
    blocks = [bspai_hip.Matrix(sb.block(s), ctx) for s in range(ns)]
    out = {"systems": ns, "rows": int(sb.matrix.n), "nnz": int(sb.matrix.nnz)}
    for prec, tol in (("fp64", 1e-8), ("mixed", 1e-4)):
        kw = dict(precision=prec, tolerance=tol, max_iterations=400)
        bspai_hip.spai_gmres_ir(mb, sb.rhs, **kw)
        t0 = time.perf_counter()
        res = bspai_hip.spai_gmres_ir(mb, sb.rhs, **kw)
        el = time.perf_counter() - t0
        out[prec] = {"batched_solves_per_s": round(ns / el, 2),
                     "succeeded": sum(r.status == "success" for r in res),
                     "spai_ms": round(res[0].spai_ms, 1), "solve_ms": round(res[0].solve_ms, 1)}
    kw = dict(precision="fp64", tolerance=1e-8, max_iterations=400)
    t0 = time.perf_counter()
    ok = 0
    for s in range(ns):
        r = bspai_hip.spai_gmres_ir(blocks[s], sb.rhs[sb.offsets[s]:sb.offsets[s + 1]], **kw)
        ok += r.status == "success"
    el = time.perf_counter() - t0
    out["fp64"]["looped_solves_per_s"] = round(ns / el, 2)
    out["fp64"]["batched_over_looped"] = round(out["fp64"]["batched_solves_per_s"] * el / ns, 2)
    return out


def bench_power_law(ctx, n):
    """Uneven SPAI columns: power-law matrix (exponent 2.2, 4..200 off-diagonals
    per row), eps 0.2, mixed IR to rel-res 1e-10; build and solve times."""
    import numpy as np

    import bspai_hip
    from bspai_hip import problems as P

    a = P.power_law(n=n)
    b = P.equal_rhs(n)
    m = bspai_hip.Matrix(a, ctx)
    log(f"power law n={n}: SPAI build (mixed, eps 0.2)")
    ctx.set_timing("spai")
    t0 = time.perf_counter()
    s = bspai_hip.construct_spai(m, 0.2, -(-n // 8), 8, precision="mixed")
    t1 = time.perf_counter()
    kt = ctx.kernel_times()
    ctx.set_timing(False)
    log(f"power law n={n}: SPAI {t1 - t0:.1f} s; mixed IR")
    r = bspai_hip.refine_with_preconditioner(m, s, b, precision="mixed", tolerance=1e-4,
                                             relres_target=1e-10)
    el = time.perf_counter() - t0
    x = np.asarray(r.solution)
    relres = float(np.linalg.norm(b - a.to_scipy() @ x) / np.linalg.norm(b))
    # per pass: kernel time and the QR rate of its accepted fits (FP32 peak
    # 157.3 TFLOP/s)
    passes = {k: {"ms": round(v["total_ms"], 1),
                  "qr_tflops": round(v["flops"] / (v["total_ms"] * 1e-3) / 1e12, 4)
                  if v["total_ms"] > 0 else 0.0}
              for k, v in kt.items()}
    return {"n": n, "nnz": int(a.nnz), "status": r.status, "relres": relres,
            "spai_ms": round(s.build_ms, 1), "passes": s.passes, "fits": s.fits,
            "pass_kernels": passes, "outer_steps": r.outer_steps,
            "inner_iterations": r.inner_iterations,
            "solve_ms": round(r.solve_ms, 1), "wall_s": round(el, 3)}


def bench_canon(ctx, cpu=True):
    """Canonical CSR from coordinate entries on the GPU (reference
    CsrMatrix::from_entries): the 128^3 system's 14,581,760 entries, each
    split into two exact halves, shuffled (29,163,520 entries), resident in
    HBM.  Algorithmic bytes: the entries read once (24 B each) and the CSR
    written once (12 B per nonzero + 4 (n+1))."""
    import numpy as np
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
    r, c, v = r[perm], c[perm], v[perm]
    dev = torch.device("cuda:0")
    rd, cd, vd = (torch.from_numpy(x).to(dev) for x in (r, c, v))
    out = bspai_hip.csr_from_entries(n, n, rd, cd, vd, ctx)
    exact = (np.array_equal(out[0].cpu().numpy(), np.asarray(a.ptr))
             and np.array_equal(out[1].cpu().numpy(), np.asarray(a.col))
             and np.array_equal(out[2].cpu().numpy().view(np.uint64), a.val.view(np.uint64)))
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        bspai_hip.csr_from_entries(n, n, rd, cd, vd, ctx)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / reps
    ctx.set_timing(True)
    bspai_hip.csr_from_entries(n, n, rd, cd, vd, ctx)
    kt = ctx.kernel_times()
    ctx.set_timing(False)
    count = r.size
    alg = 24.0 * count + 12.0 * a.nnz + 4.0 * (n + 1)
    line = {"entries": int(count), "n": n, "nnz": int(a.nnz), "equals_generator_csr": exact,
            "ms_per_call": round(ms, 3), "entries_per_s": round(count / (ms * 1e-3), 1),
            "algorithmic_bytes": alg,
            "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1),
                         "peak": 8000.0, "unit": "GB/s",
                         "frac": round(alg / (ms * 1e-3) / 8e12, 4)},
            "kernels": {k: round(x["avg_us"], 1) for k, x in kt.items()}}
    if cpu:
        sys.path.insert(0, REPO)
        from oracle import pyoracle as O

        t0 = time.perf_counter()
        O.from_entries(n, n, r, c, v)
        el = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(count / el, 1), "unit": "entries/s", "cores": 1,
                                "kind": "port", "sample": "oracle from_entries on the same "
                                "29,163,520 entries (std::stable_sort, one core)",
                                "ms": round(1e3 * el, 1)}
    return line


PEAK_TFLOPS = {"double": 78.6, "float": 157.3}


def kernel_rooflines(kt):
    """Roofline rows from the library's per-kernel event totals, largest share
    first: HBM-bound kernels report algorithmic GB/s against 8 TB/s; the SPAI
    column kernels (dense QR on the vector units) report algorithmic TFLOP/s
    against the FP64 or FP32 peak."""
    total = sum(v["total_ms"] for v in kt.values()) or 1.0
    rows = []
    for name, v in kt.items():
        sec = v["total_ms"] * 1e-3
        if name.startswith(("spai_columns_kernel", "spai_wave_kernel", "spai_block_kernel")):
            prec = "float" if "<float>" in name else "double"
            ach = v["flops"] / sec / 1e12 if sec > 0 else 0.0
            # vector-ALU code (no MFMA instruction): priced against the
            # chip's FP32 / FP64 peak, which the vector units share
            bound, unit, peak = "valu", "TFLOP/s", PEAK_TFLOPS[prec]
            per = {"flops": v["flops"] / max(1, v["launches"])}
        else:
            ach = v["bytes"] / sec / 1e9 if sec > 0 else 0.0
            # kernels with no algorithmic bytes (the Gram solve: one wave's
            # serial chain over a few KB) are latency-bound, not HBM-bound
            bound = "hbm" if v["bytes"] > 0 else "latency"
            unit, peak = "GB/s", HBM_PEAK_GBS
            per = {"bytes": v["bytes"] / max(1, v["launches"])}
        rows.append(dict(kernel=name, launches=v["launches"], total_ms=round(v["total_ms"], 3),
                         avg_us=round(v["avg_us"], 2), share=round(v["total_ms"] / total, 4),
                         bound=bound, achieved=round(ach, 3), peak=peak, unit=unit,
                         frac=round(ach / peak, 4), per_launch=per))
    rows.sort(key=lambda r: -r["total_ms"])
    return rows


def log(msg):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def compare_rows(p_a, p_b):
    """Rows of two preconditioners P (rows of P are the columns of M): the
    number whose patterns differ, and the largest per-row normwise relative
    difference max|a - b| / max|b| over the rows whose patterns agree."""
    import numpy as np

    ptr_a, col_a, val_a = (np.asarray(x) for x in p_a)
    ptr_b, col_b, val_b = (np.asarray(x) for x in p_b)
    n = len(ptr_b) - 1
    la, lb = np.diff(ptr_a), np.diff(ptr_b)
    rows = np.nonzero((la == lb) & (lb > 0))[0]
    L = lb[rows]
    offs = np.arange(L.sum()) - np.repeat(np.cumsum(L) - L, L)
    ia = np.repeat(ptr_a[:-1][rows], L) + offs
    ib = np.repeat(ptr_b[:-1][rows], L) + offs
    starts = np.cumsum(L) - L
    same = np.add.reduceat((col_a[ia] != col_b[ib]).astype(np.int64), starts) == 0
    num = np.maximum.reduceat(np.abs(val_a[ia] - val_b[ib]), starts)
    den = np.maximum.reduceat(np.abs(val_b[ib]), starts)
    rel = (num / den)[same]
    return int(n - same.sum()), float(rel.max()) if rel.size else 0.0


def cpu_baseline(a, b_host, p64, p32, solutions):
    """One complete solve of the benchmark system by the oracle (the CPU FP64
    restatement of the reference's cpu_fp64_ir path, bit-identical to it) on
    one host core, timed with perf_counter around the public calls: SPAI
    construction, then refinement with that preconditioner (outercap 20,
    innercap 200, tau 1e-8, the reference's stop).  The GPU results are then
    checked against it: FP64 P bits, FP32 P pattern differences and column
    errors, and both solutions."""
    import numpy as np

    sys.path.insert(0, REPO)
    from oracle import pyoracle as O
    from bspai_hip import problems as P

    n = a.n
    O.set_progress(lambda phase, done, value: log(
        f"cpu baseline: SPAI columns {done}/{value}" if phase == 0 else
        f"cpu baseline: refinement step {done} done ({value} inner iterations)"))
    log(f"cpu baseline: oracle SPAI on all {n} columns, one core")
    t0 = time.perf_counter()
    s = O.construct_spai(a, 0.3, -(-n // 8), 8)
    t_spai = time.perf_counter() - t0
    log(f"cpu baseline: SPAI {t_spai:.1f} s ({s.status}); refinement")
    t0 = time.perf_counter()
    r = O.spai_gmres_ir(a, b_host, 0.3, tolerance=1e-8, max_iterations=200, max_outer_steps=20,
                        p=P.Csr(n, s.p_ptr, s.p_col, s.p_val))
    t_ir = time.perf_counter() - t0
    O.set_progress(None)
    log(f"cpu baseline: refinement {t_ir:.1f} s ({r.status}, inner {r.inner_iterations})")
    total = t_spai + t_ir
    x_ref = np.asarray(r.solution)
    relres = float(np.linalg.norm(b_host - a.to_scipy() @ x_ref) / np.linalg.norm(b_host))
    ref_p = (s.p_ptr, s.p_col, s.p_val)
    nd32, err32 = compare_rows(p32, ref_p)
    check = {
        "oracle": {"status": r.status, "outer_steps": r.outer_steps,
                   "inner_iterations": r.inner_iterations, "relres": relres,
                   "nbe": float(r.history[-1][2]) if len(r.history) else None},
        "fp64_spai_bits_equal_oracle": bool(
            np.array_equal(np.asarray(p64[0]), s.p_ptr) and np.array_equal(np.asarray(p64[1]), s.p_col)
            and np.array_equal(np.asarray(p64[2]).view(np.uint64), s.p_val.view(np.uint64))),
        "fp32_spai_vs_oracle": {"pattern_differing_columns": nd32,
                                "max_column_rel_error_on_common_patterns": err32},
        "solution_max_rel_diff_vs_oracle": {
            k: float(np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref)))
            for k, x in solutions.items()},
    }
    base = {"value": round(1.0 / total, 6), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"one complete solve of the same {n}-row system by the oracle (CPU FP64 "
                      "restatement of the reference's cpu_fp64_ir path) on one host core, timed: "
                      f"SPAI {t_spai:.1f} s + refinement {t_ir:.1f} s (inner "
                      f"{r.inner_iterations}); the reference binary's own SPAI scans every row "
                      "per fit (O(n^2)) and cannot run at this size",
            "seconds": round(total, 3), "spai_s": round(t_spai, 3), "ir_s": round(t_ir, 3),
            "host_cpus": os.cpu_count()}
    return base, check


def cpu_baseline_c1(trials):
    """The reference itself (its CPU sources compiled into oracle/_ref/ref_time)
    on the CPU-runnable configuration, the 2-D 64x64 system (eps 0.3, outercap
    20, innercap n), timed around spai_gmres_ir after `trials` warm-ups; the
    oracle port when the compiled reference is absent."""
    import json as _json
    import subprocess
    import tempfile

    import numpy as np

    from bspai_hip import problems as P

    a = P.convection_diffusion_2d()
    exe = os.path.join(REPO, "oracle", "_ref", "ref_time")
    if os.path.exists(exe):
        log(f"cpu baseline (64x64): the reference, {trials} + {trials} trials")
        with tempfile.TemporaryDirectory() as d:
            mtx = os.path.join(d, "c1.mtx")
            P.write_matrix_market(mtx, a)
            out = subprocess.run([exe, mtx, "0.3", "20", "0", str(trials), str(trials)],
                                 capture_output=True, text=True, check=True, timeout=600)
        res = _json.loads(out.stdout)
        kind, secs = "reference", res["seconds"]
        inner = res["inner"]
    else:
        sys.path.insert(0, REPO)
        from oracle import pyoracle as O

        log(f"cpu baseline (64x64): the oracle port, {trials} + {trials} trials")
        b = P.equal_rhs(a.n)
        secs = []
        for t in range(2 * trials):
            t0 = time.perf_counter()
            r = O.spai_gmres_ir(a, b, 0.3)
            if t >= trials:
                secs.append(time.perf_counter() - t0)
        kind, inner = "port", r.inner_iterations
    med = float(np.median(secs))
    return {"value": round(1.0 / med, 4), "unit": "solves/s", "cores": 1, "kind": kind,
            "sample": f"2-D 64x64 convection-diffusion (n=4096), eps 0.3, outercap 20, innercap n, "
                      f"{trials} warm-up + {trials} measured trials, median; inner {inner}",
            "seconds": [round(x, 4) for x in secs]}


if __name__ == "__main__":
    main()
