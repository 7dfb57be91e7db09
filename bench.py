#!/usr/bin/env python3
"""Benchmark of the multifrontal factorization on one MI355X.

One step = one complete factorization of the synthetic workload (extend-add
assembly of every front + partial factorization of every front, all tree
levels, leaves first) with inputs already resident in HBM.  Reported value:
algorithmic factor TFLOP/s (LAWN-41 count of the reference's
Mesh_::estimate_cost, the same numerator whatever precision executes it) of
the mixed-precision path (FP32 fronts and FP32 MFMA Schur updates below the
top tree levels, FP64 top-level pivot blocks), provided its FP64-residual
refinement reaches backward error 1e-12; the plain FP64 path is measured on
the same matrices beside it.

Default workload (the configuration the target is stated on): 3D unit cube,
16^3 hexahedral Q4 elements (274,625 dofs, 8191 fronts, root separator
4225), SPD element matrices K_e = (M^T M + 125 I) 10^u with u ~ U(-2, 2),
Cholesky, one right-hand side.  `--workload lu` runs the 3D Q3 LU
configuration (117,649 dofs, K_e = (M^T M + 64 I + 0.1 S) lognormal(0, 1)).

Printed: one JSON line with the metric, both paths' factor throughput, the
time from the start of factorization to the first iterate with backward
error <= 1e-12 for each path and their ratio, the live roofline of the
dominant kernel (the Schur-update GEMM) and the CPU oracle timed on a
bounded sample of the same tree.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "factor TFLOP/s + time-to-1e-12 backward error, mixed-prec vs FP64 GPU"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 MFMA
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 MFMA

WORKLOADS = {
    "chol": dict(kind=2, dims=(16, 16, 16), p=4, method=1, alpha=125.0, beta=0.0, scale_kind=2, p0=-2.0, p1=2.0,
                 name="3D 16^3 hex Q4, SPD K_e=(M^T M+125 I)*10^U(-2,2), Cholesky, 1 RHS"),
    "lu": dict(kind=2, dims=(16, 16, 16), p=3, method=3, alpha=64.0, beta=0.1, scale_kind=1, p0=1.0, p1=0.0,
               name="3D 16^3 hex Q3, K_e=(M^T M+64 I+0.1 S)*lognormal(0,1), LU partial pivoting, 1 RHS"),
    # batched uneven fronts: FP32 fronts, FP64 refinement, 64 RHS, graph-replayed
    # re-factorization over value reseeds on the fixed pattern
    "varp": dict(kind=3, dims=(256, 256), p=8, mesh_seed=12345, method=3, alpha=-1.0, beta=0.1, scale_kind=0,
                 p0=0.0, p1=0.0, nrhs=64, top=0, reseeds=10, refine="auto",
                 name="2D 256^2 quads, p_e ~ U{1..8} (min-p edges), K_e=M^T M+n I+0.1 S, LU, FP32 fronts + FP64 "
                      "refinement, 64 RHS, graph-replayed re-factorization over 10 reseeds"),
    # low precision: Cholesky, 256 RHS
    "q3big": dict(kind=2, dims=(24, 24, 24), p=3, method=1, alpha=64.0, beta=0.0, scale_kind=1, p0=0.5, p1=0.0,
                  nrhs=256, top=1, refine="auto",
                  name="3D 24^3 hex Q3, SPD K_e=(M^T M+64 I)*lognormal(0,0.5), Cholesky, 256 RHS"),
}


REFINE_KINDS = {"classical": lambda U: U.REFINE_CLASSICAL, "gmres": lambda U: U.REFINE_GMRES,
                "auto": lambda U: U.REFINE_AUTO}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_sync():
    try:
        ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()
    except OSError:
        pass


def build_mesh(U, w, seed):
    m = U.Mesh()
    m.build_synthetic(w["kind"], w["dims"], w["p"], w.get("mesh_seed", 0))
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, w.get("nrhs", 1))
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=w["alpha"], beta=w["beta"], scale_kind=w["scale_kind"], p0=w["p0"],
                        p1=w["p1"])
    m.random_matrix()
    m.set_rhs()
    return m


def cpu_baseline(U, m, w, repeats=3):
    """The CPU oracle (the C restatement of the reference algorithm in
    oracle/: BLAS-free blocked loops, OpenMP over the fronts of a level, or
    inside a front for levels with fewer fronts than threads) factorizing
    the whole tree of the same matrices the GPU factorized (leaf stages and
    right-hand sides fetched from the device), tile b = 256; the best of
    `repeats` runs, then one solve, whose solution is returned for the
    forward errors of the GPU paths."""
    from oracle.oracle import OracleMesh, set_threads

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    threads = max(1, min(threads, 16, os.cpu_count() or 1))
    set_threads(threads)
    om = OracleMesh.from_connectivity(m.export_connectivity())
    om.init_numeric(m.n_rhs)
    for f in range(m.n_fronts()):
        fi = m.front_info(f)
        if fi["leaf"]:
            om.set_front(fi["id"], m.fetch_stage(f))
            om.set_rhs_front(fi["id"], m.fetch_rhs_front(f))
    flops = om.flops(w["method"])
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        info = om.factor(w["method"], 256)
        times.append(time.perf_counter() - t0)
    om.solve(w["method"])
    x = om.x_global(m.get_n_dof())
    dt = min(times)
    out = dict(value=round(flops / dt / 1e12, 5), unit="TFLOP/s", cores=threads, kind="port",
               sample="BLAS-free oracle (C restatement of the reference's lu_piv_hier / chol_hier, blocked loops, "
                      "OpenMP) factorizing the whole tree (%d fronts, %.1f GFLOP) of the same device-generated "
                      "matrices, tile b=256, %d threads, best of %d: %s s, info %d"
                      % (m.n_fronts(), flops / 1e9, threads, repeats, ", ".join("%.2f" % t for t in times), info))
    return out, x


def run_path(U, m, w, mode, top, args):
    """Factor throughput, time-to-1e-12 and live kernel stats of one precision
    path on the already generated matrices."""
    chol = w["method"] == 1
    factor = m.chol_with_free if chol else m.lu_piv_with_free
    solve = m.solve_chol if chol else m.solve_lu_piv
    name = {U.PREC_FP64: "fp64", U.PREC_MIXED: "mixed", U.PREC_MIXED_BF16: "bf16", U.PREC_MIXED_FP16: "fp16"}[mode]
    m.set_precision(mode, top)
    m.set_refinement(REFINE_KINDS[args.refine or w.get("refine", "classical")](U), 20, 1e-6)
    tuned = None
    if args.tune:
        # the project's autotuner picks the schedule knobs on this pattern
        t_ms, tuned = m.autotune(w["method"])
        log("[%s] autotune: %.2f ms with %s" % (name, t_ms, tuned))
    for i in range(args.warmup):
        info = factor()
        log("[%s] warmup %d: %.2f ms (info %d)" % (name, i, m.timing()[0], info))
    device_sync()
    t0 = time.perf_counter()
    dev_ms = []
    for i in range(args.steps):
        info = factor()
        dev_ms.append(m.timing()[0])
    device_sync()
    ms_step = (time.perf_counter() - t0) / args.steps * 1e3
    log("[%s] timed %d steps: %.3f ms/step wall, device %.3f ms/step"
        % (name, args.steps, ms_step, sum(dev_ms) / len(dev_ms)))
    if info:
        raise SystemExit("zero / non-positive pivot in element %d" % (info - 1))
    # time from the start of factorization to the first iterate with
    # backward error <= 1e-12 (refinement stops there; the FP64 path needs
    # no correction, only the residual that verifies it)
    solve()
    m.improve(w["method"], 1e-12, 30)  # assembles the CSR once (host symbolic)
    device_sync()
    t0 = time.perf_counter()
    factor()
    solve()
    n_corr = m.improve(w["method"], 1e-12, 30)
    device_sync()
    t_sol = (time.perf_counter() - t0) * 1e3
    hist, ref_ms = m.refine_history()
    n_solves = m.refine_solves()
    eta = m.backward_error()
    x = m.fetch_x()
    f_ms, s_ms = m.timing()
    log("[%s] factor+solve+refine %.2f ms (factor %.2f, solve %.2f, refine %.2f ms, %d corrections, %d solves), "
        "backward error %s" % (name, t_sol, f_ms, s_ms, ref_ms, n_corr, n_solves,
                               " -> ".join("%.2e" % e for e in hist)))
    m.set_profile(True)
    factor()
    stats = m.kernel_stats()
    m.set_profile(False)
    for k, (ms_k, fl_k, by_k, n_k) in stats.items():
        if n_k:
            log("  %-10s %5d launches %9.3f ms  %8.3f TFLOP/s  %8.1f GB/s"
                % (k, n_k, ms_k, fl_k / ms_k / 1e9 if ms_k else 0, by_k / ms_k / 1e6 if ms_k else 0))
    return dict(ms_step=ms_step, t_sol=t_sol, eta=eta, hist=[float(e) for e in hist], n_corr=n_corr, solve_ms=s_ms,
                refine_ms=ref_ms, n_solves=n_solves, stats=stats, tuned=tuned, x=x)


def run_reseeds(U, m, w, args, n):
    """Re-factorization of new element values on the fixed pattern: the
    factorization and solves are graph-captured once and replayed for each
    reseed (device-generated values), each refined to backward error 1e-12."""
    chol = w["method"] == 1
    factor = m.chol_with_free if chol else m.lu_piv_with_free
    solve = m.solve_chol if chol else m.solve_lu_piv
    m.set_precision(U.PREC_MIXED, w.get("top", 0))
    m.set_refinement(REFINE_KINDS[args.refine or w.get("refine", "classical")](U), 20, 1e-6)
    m.set_graph(True)
    out = []
    for seed in range(1, n + 1):
        m.set_random_params(seed=seed, alpha=w["alpha"], beta=w["beta"], scale_kind=w["scale_kind"], p0=w["p0"],
                            p1=w["p1"])
        m.random_matrix()
        m.set_rhs()
        device_sync()
        t0 = time.perf_counter()
        info = factor()
        solve()
        n = m.improve(w["method"], 1e-12, 30)
        device_sync()
        t = (time.perf_counter() - t0) * 1e3
        f_ms, s_ms = m.timing()
        hist, ref_ms = m.refine_history()
        out.append(dict(seed=seed, info=info, ms=t, factor_ms=f_ms, solve_ms=s_ms, refine_ms=ref_ms, corrections=n,
                        solves=m.refine_solves(), eta=float(m.backward_error())))
        log("[reseed %d] %.2f ms (factor %.2f, solve %.2f, refine %.2f, %d corrections, %d solves) eta %.2e info %d"
            % (seed, t, f_ms, s_ms, ref_ms, n, m.refine_solves(), m.backward_error(), info))
    m.set_graph(False)
    return out


# HBM bytes per k_gemm launch (the gemm + schur launches: pure products, not
# the panel updates that also factor diagonal blocks) from rocprofv3
# FETCH_SIZE / WRITE_SIZE passes over `bench.py --traffic-probe`, calibrated
# against a copy of known size
# (profiles/k_gemm_chol_q4_16cube_hbm_traffic.md): mixed path of the chol
# workload only; other paths and workloads report null.
GEMM_TRAFFIC = {("chol", "mixed"): 738.0e6}


def gemm_roofline(stats, peak, kernel, traffic=None):
    g_ms = stats["gemm"][0] + stats["schur"][0]
    g_fl = stats["gemm"][1] + stats["schur"][1]
    g_n = stats["gemm"][3] + stats["schur"][3]
    achieved = g_fl / (g_ms * 1e-3) / 1e12 if g_ms else 0.0
    return {"kernel": kernel, "bound": "mfma", "achieved": round(achieved, 4), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic, "launches": g_n,
            "avg_launch_us": round(g_ms / g_n * 1e3, 3) if g_n else None}


PEAK_HBM_GBS = 8000.0  # MI355X HBM3E
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 / FP16 MFMA


def lowprec_roofline(stats, kernel):
    """The final Schur complement products on packed BF16 / FP16 operands
    (k_schur_lp launches of precision modes 2 / 3) against the dense BF16 /
    FP16 MFMA peak."""
    ms, fl, _, n = stats["schur_lp"]
    if not n:
        return None
    achieved = fl / (ms * 1e-3) / 1e12
    return {"kernel": kernel, "bound": "mfma", "achieved": round(achieved, 4), "peak": PEAK_BF16_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16_TFLOPS, 4), "traffic": None, "launches": n,
            "avg_launch_us": round(ms / n * 1e3, 3)}


def assemble_roofline(stats):
    """Extend-add (k_assemble) against HBM bandwidth: algorithmic bytes
    (every stored parent entry written once, its own stage and each child's
    ABR entry read once) over the summed launch durations."""
    ms, _, by, n = stats["assemble"]
    achieved = by / (ms * 1e-3) / 1e9 if ms else 0.0
    return {"kernel": "k_assemble (extend-add, gather form)", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
            "launches": n, "avg_launch_us": round(ms / n * 1e3, 3) if n else None}


def traffic_probe(U, m, w, top):
    """The program of the HBM-traffic counter passes: a device-to-device
    copy of 512 MiB (the runtime's copy kernel, 512 MiB read and 512 MiB
    written) whose counter readings calibrate FETCH_SIZE / WRITE_SIZE, then
    two mixed-path factorizations; the second one's k_gemm dispatches are
    the measured ones."""
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = ctypes.c_size_t(512 << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), nbytes) != 0:
            raise SystemExit("hipMalloc failed")
    if hip.hipMemcpy(dst, src, nbytes, 3) != 0 or hip.hipDeviceSynchronize() != 0:  # device to device
        raise SystemExit("calibration copy failed")
    hip.hipFree(src)
    hip.hipFree(dst)
    m.set_precision(U.PREC_MIXED, top)
    factor = m.chol_with_free if w["method"] == 1 else m.lu_piv_with_free
    for _ in range(2):
        info = factor()
    device_sync()
    log("traffic probe: 2 factorizations, info %d" % info)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="chol")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the compact lines of the other configurations (16^3 Q3 LU, 256^2 var-p, 24^3 Q3)")
    ap.add_argument("--nb", type=int, default=64, help="panel width")
    ap.add_argument("--top", type=int, default=2, help="tree levels kept in FP64 by the mixed path")
    ap.add_argument("--refine", choices=sorted(REFINE_KINDS), default=None,
                    help="refinement kind (default: the workload's; classical for chol / lu)")
    ap.add_argument("--reseeds", type=int, default=None,
                    help="graph-replayed re-factorizations after the timed paths (default: the workload's)")
    ap.add_argument("--bf16", action="store_true",
                    help="also time the BF16-Schur mixed path (BF16 operands, FP32 accumulation in the Schur "
                         "complement products), reported beside the FP32 one (always on for the default "
                         "workload)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="counter-pass program instead of the bench: a copy kernel of known size (calibration: "
                         "512 MiB read, 512 MiB written), then two factorizations of the mixed path (the second "
                         "is the measured one); run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="schedule knob (uhm_set_tuning_) applied before timing, repeatable")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner over the schedule knobs before timing (default: built-in knobs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one process, one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import uhm_amd as U
    U.set_hier_block_size(args.nb)
    if args.traffic_probe:
        w = WORKLOADS[args.workload]
        m = build_mesh(U, w, seed=3)
        traffic_probe(U, m, w, w.get("top", args.top))
        return

    # the headline workload in full, with the CPU baseline on its matrices
    res = measure(U, args.workload, args, args.steps, args.warmup, bf16=args.bf16 or args.workload == "chol")
    w, m, flops = res["w"], res["m"], res["flops"]
    r64, rmx, rbf = res["fp64"], res["mixed"], res["bf16"]
    mixed_ok = rmx["eta"] <= 1e-12
    head = rmx if mixed_ok else r64
    cpu, x_oracle = None, None
    if not args.no_cpu:
        try:
            cpu, x_oracle = cpu_baseline(U, m, w)
            log("cpu baseline:", cpu["value"], cpu["unit"], cpu["sample"])
        except Exception as e:  # the baseline is reported, never the product
            log("cpu baseline failed:", repr(e))
    top = w.get("top", args.top)
    out = {
        "metric": METRIC,
        "value": _tflops(flops, head),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fronts + f64 top levels and refinement" if mixed_ok else "f64",
        "data": "synthetic (Philox element matrices generated on the device)",
        "config": {
            "workload": w["name"],
            "ndof": m.get_n_dof(),
            "fronts": m.n_fronts(),
            "factor_gflop": round(flops / 1e9, 3),
            "precision_path": ("mixed: FP32 fronts and FP32 MFMA Schur updates below the top %d tree levels, "
                               "FP64 top-level pivot blocks, FP64-residual refinement to backward error 1e-12"
                               % top) if mixed_ok else "fp64 (mixed path did not reach 1e-12)",
            "refinement": args.refine or w.get("refine", "classical"),
            "panel_width": args.nb,
            "n_rhs": w.get("nrhs", 1),
        },
        "time_to_1e-12_ms": _ttt(res),
        "roofline": gemm_roofline(head["stats"], PEAK_FP32_TFLOPS if mixed_ok else PEAK_FP64_TFLOPS,
                                  "k_gemm (Schur / trailing update, %s MFMA)" % ("FP32" if mixed_ok else "FP64"),
                                  GEMM_TRAFFIC.get((args.workload, "mixed" if mixed_ok else "fp64"))),
        "roofline_fp64_path": gemm_roofline(r64["stats"], PEAK_FP64_TFLOPS, "k_gemm (FP64 MFMA)"),
        "roofline_extend_add": assemble_roofline(head["stats"]),
        "roofline_schur_bf16": lowprec_roofline(rbf["stats"], "k_schur_lp (packed BF16 operands)") if rbf else None,
        "roofline_schur_fp16": (lowprec_roofline(res["fp16"]["stats"], "k_schur_lp (packed FP16 operands)")
                                if res["fp16"] else None),
        "mixed": _path(flops, rmx, r64, x_oracle),
        "fp64": _path(flops, r64, r64, x_oracle),
        "mixed_bf16": _path(flops, rbf, r64, x_oracle) if rbf else None,
        "mixed_fp16": _path(flops, res["fp16"], r64, x_oracle) if res["fp16"] else None,
        "reseeds": res["reseeds"],
        "cpu_baseline": cpu,
    }
    del m, res
    # the other configurations, compact (driver-timed in the same run)
    extra = {}
    if args.workload == "chol" and not args.no_extra:
        for name in ("lu", "varp", "q3big"):
            try:
                rx = measure(U, name, args, EXTRA_STEPS, 1, bf16=(name == "q3big"))
                extra[name] = _compact(rx)
                del rx
            except Exception as e:  # reported, the headline stands
                log("config %s failed: %r" % (name, e))
                extra[name] = {"error": repr(e)}
    out["configs"] = extra or None
    print(json.dumps(out), flush=True)


EXTRA_STEPS = 3


def measure(U, name, args, steps, warmup, bf16=False):
    """Mesh, FP64 and mixed paths (and the BF16-Schur path if asked) of one
    workload, plus its graph-replayed reseeds."""
    import copy
    w = WORKLOADS[name]
    a = copy.copy(args)
    a.steps, a.warmup = steps, warmup
    t0 = time.perf_counter()
    m = build_mesh(U, w, seed=3)
    flops, _, _, _ = m.estimate_cost(w["method"])
    log("mesh %s: ndof %d, fronts %d, %.1f GFLOP, setup %.1f s"
        % (name, m.get_n_dof(), m.n_fronts(), flops / 1e9, time.perf_counter() - t0))
    for kv in args.set:
        k, v = kv.split("=")
        m.set_tuning(int(k), int(v))
    top = w.get("top", args.top)
    r64 = run_path(U, m, w, U.PREC_FP64, top, a)
    rmx = run_path(U, m, w, U.PREC_MIXED, top, a)
    rbf = run_path(U, m, w, U.PREC_MIXED_BF16, top, a) if bf16 else None
    rfp = run_path(U, m, w, U.PREC_MIXED_FP16, top, a) if bf16 else None
    n_reseeds = w.get("reseeds", 0) if args.reseeds is None else args.reseeds
    reseeds = run_reseeds(U, m, w, a, n_reseeds) if n_reseeds else None
    return dict(name=name, w=w, m=m, flops=flops, fp64=r64, mixed=rmx, bf16=rbf, fp16=rfp, reseeds=reseeds)


def _rel(a, b):
    import numpy as np
    return float(np.abs(a - b).max() / np.abs(b).max())


def _tflops(flops, r):
    return round(flops / (r["ms_step"] * 1e-3) / 1e12, 4)


def _ttt(res):
    """Time from the start of factorization to backward error <= 1e-12 for
    each path, and the FP64 path's time over each low-precision path's."""
    r64 = res["fp64"]
    out = {"fp64": round(r64["t_sol"], 3) if r64["eta"] <= 1e-12 else None}
    for key, r in (("mixed", res["mixed"]), ("mixed_bf16", res["bf16"]), ("mixed_fp16", res["fp16"])):
        if r is None:
            continue
        ok = r["eta"] <= 1e-12
        out[key] = round(r["t_sol"], 3) if ok else None
        out["speedup" + key[5:]] = round(r64["t_sol"] / r["t_sol"], 3) if ok else None
    return out


def _path(flops, r, r64, x_oracle):
    """One path's record: factor throughput, time to 1e-12, backward error
    history, forward errors against the FP64 path and the oracle, per-kind
    kernel statistics."""
    return {"factor_tflops": _tflops(flops, r), "factor_ms": round(r["ms_step"], 4),
            "time_to_1e-12_ms": round(r["t_sol"], 3) if r["eta"] <= 1e-12 else None,
            "backward_error": r["eta"], "backward_error_history": r["hist"],
            "forward_error_vs_fp64": _rel(r["x"], r64["x"]) if r is not r64 else 0.0,
            "forward_error_vs_oracle": _rel(r["x"], x_oracle) if x_oracle is not None else None,
            "corrections": r["n_corr"], "refine_solves": r["n_solves"],
            "solve_ms": round(r["solve_ms"], 3), "refine_ms": round(r["refine_ms"], 3), "schedule": r["tuned"],
            "kernels": {k: {"ms": round(v[0], 4), "gflop": round(v[1] / 1e9, 3),
                            "gbytes": round(v[2] / 1e9, 4), "launches": v[3]}
                        for k, v in r["stats"].items() if v[3]}}


def _compact(res):
    """Short record of a secondary configuration."""
    w, flops, r64 = res["w"], res["flops"], res["fp64"]
    out = {"workload": w["name"], "ndof": res["m"].get_n_dof(), "fronts": res["m"].n_fronts(),
           "factor_gflop": round(flops / 1e9, 3), "n_rhs": w.get("nrhs", 1), "steps": EXTRA_STEPS,
           "time_to_1e-12_ms": _ttt(res)}
    for key, r in (("fp64", r64), ("mixed", res["mixed"]), ("mixed_bf16", res["bf16"]), ("mixed_fp16", res["fp16"])):
        if r is None:
            continue
        out[key] = {"factor_ms": round(r["ms_step"], 4), "factor_tflops": _tflops(flops, r),
                    "solve_ms": round(r["solve_ms"], 3), "refine_ms": round(r["refine_ms"], 3),
                    "corrections": r["n_corr"], "refine_solves": r["n_solves"], "backward_error": r["eta"],
                    "forward_error_vs_fp64": _rel(r["x"], r64["x"]) if r is not r64 else 0.0,
                    "roofline_gemm": gemm_roofline(r["stats"], PEAK_FP64_TFLOPS if key == "fp64" else PEAK_FP32_TFLOPS,
                                                   "k_gemm"),
                    "roofline_extend_add": assemble_roofline(r["stats"])}
    if res["reseeds"]:
        rs = res["reseeds"]
        out["reseeds"] = {"count": len(rs), "max_eta": max(r["eta"] for r in rs),
                          "mean_ms": round(sum(r["ms"] for r in rs) / len(rs), 3),
                          "corrections": [r["corrections"] for r in rs]}
    return out


if __name__ == "__main__":
    main()
