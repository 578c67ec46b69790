"""Benchmark of the per-Newton-step mixed-FEM hot path on one MI355X.

Workload (N=1): the precision config of the benchmark set -- a jittered
100x40x40x6 Kuhn beam (960,000 tets), per-tet Young's modulus log-uniform in
[1e5, 1e6], nu = 0.45, Stable Neo-Hookean, mixed-stretch variables, BDF1
dt = 1/30, gravity, clamped end, Jacobi CG with the reference defaults
(tolerance 1e-4, 500 iterations, 5 Newton iterations). A bench "step" is one
timestep (NewtonOptimizer::step); the metric counts Newton iterations x tets
per second over the timed timesteps.

The mixed FP32/FP64 mode is the reported `value`; the FP64 device mode runs
the same timesteps for the comparison. A second FP64 run with another CG
kernel shape (dot products summed in another order) measures how far
rounding alone moves the trajectory. The run is marked invalid (exit status 3
after the JSON line) when the mixed positions leave the FP64 ones by more
than 1e-6 relative (2-norm), or by more than 3x that FP64 rounding spread
where the spread is the larger, or when mixed mode fails more line searches
than FP64 (not counting those at Newton iterations that have reached FP64
rounding, where the energy test is chance in either precision). The FP64 CPU oracle is timed on a bounded sample of the
same scene on the host cores: one warm-up timestep, then the median of three
timesteps.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cg_bytes(nfree, nnzb, vb):
    """Algorithmic bytes of one CG iteration (one k_cg_iter launch): the BSR
    3x3 matrix (values + block column ids + row pointers) read once, the
    per-row state {r, q, p, D^-1} read once and written once, and the FP64
    solution read and written once."""
    n = 3 * nfree
    return nnzb * (9 * vb + 4) + 4 * (nfree + 1) + n * (2 * 4 * vb + 16)


def minres_tet_bytes(nT, nV, nfree, vb):
    """Algorithmic bytes of one k_mr_tet launch (one MINRES iteration's element
    kernel): per tet the update of the previous iteration on its 12 element
    entries (w, p_{k-2}, p_{k-3}, x read; p_{k-1}, x written), the product
    (w read, v_old read and written, the 12-entry x-row contribution written,
    ds/dF (54 values), h^2 Hessian (21), shape gradients (12 FP64), volume,
    corner ids), plus the x part of w (3 per free vertex) and the free map
    gathered once."""
    per_tet = 6 * 96 + 96 + 2 * 96 + 96 + 75 * vb + 96 + 8 + 16
    return nT * per_tet + nfree * 24 + nV * 4


def measured_traffic(lanes, unroll, workgroup, default_workload):
    """HBM bytes per CG-kernel launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (copy-calibrated) on the default workload, for the shape the
    tuner picked; None for other workloads."""
    path = os.path.join(ROOT, "profiles", "k_cg_iter_c3_mixed_traffic.json")
    if not default_workload or not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    per = d["per_shape_bytes"]
    key = f"k_cg_iter<{lanes}, {unroll}, {workgroup}, float>"
    vals = per.get(key) or sorted(per.values(), key=lambda v: v["read"] + v["write"])[len(per) // 2]
    return vals["read"] + vals["write"]


def run_gpu(scene, precision, steps, warmup, profile=True, element_kernel=0):
    from mfem_amd import native
    sim = native.Sim(scene, precision=precision)
    if element_kernel:
        sim.set_element_kernel(element_kernel)
    for _ in range(warmup):
        sim.step()
    if profile:
        sim.set_profiling(1)
    newton = pcg = 0
    spmv_ms = 0.0
    spmv_n = 0
    t0 = time.perf_counter()
    per_step = []
    ls_fail = []
    floor_fail = []
    halvings = []
    for k in range(steps):
        ts = time.perf_counter()
        sim.step()
        st = sim.stats()
        per_step.append(time.perf_counter() - ts)
        hist = sim.newton_history()
        halv = int(hist[:, 9].sum()) if hist.size else 0
        ls_fail.append(int(st["ls_fail"]))
        # line searches that ran out of halvings where the Newton iteration
        # had reached FP64 rounding (energy change < 1e-10 relative, max |dx|,
        # |ds| < 1e-6): E(alpha) < E(0) is chance there in either precision,
        # and the skipped step moves no position by more than 1e-6
        floor_fail.append(int(np.sum((hist[:, 9] >= scene.ls_iters) & (hist[:, 2] < 1e-10)
                                     & (hist[:, 3] < 1e-6))) if hist.size else 0)
        halvings.append(halv)
        print(f"[bench] {precision} step {k}: {per_step[-1] * 1e3:.1f} ms, newton {st['newton_iters']}, "
              f"cg {st['pcg_iters']}, sweeps {st['refinements']}, ls halvings {halv}, "
              f"ls failures {st['ls_fail']}", file=sys.stderr, flush=True)
        newton += st["newton_iters"]
        pcg += st["pcg_iters"]
        spmv_ms += st["spmv_ms"]
        spmv_n += st["spmv_count"]
    elapsed = time.perf_counter() - t0
    nfree, nnzb, vb = sim.matrix_info()
    x = sim.positions()
    # per-phase device times (update, solve, line search) from one more,
    # untimed timestep with the phase events on: each event is a marker in
    # the stream that costs the device a few microseconds
    sim.set_profiling(3 if profile else 2)
    sim.step()
    st = sim.stats()
    tuning = None
    if scene.solver == "eigen-pcg-diag":
        trials, chosen, shape = sim.cg_tuning()
        spmv_alone_ms = sim.time_spmv(50)
        tuning = {"trials_variant_grid_ms": [[v, g, round(t, 5)] for v, g, t in trials],
                  "chosen": chosen, "lanes_per_row": shape[0], "blocks_in_flight": shape[1],
                  "workgroup": shape[2], "grid": shape[3], "plain_spmv_ms": spmv_alone_ms}
    sim.close()
    return dict(elapsed=elapsed, newton=newton, pcg=pcg, spmv_ms=spmv_ms, spmv_n=spmv_n,
                nfree=nfree, nnzb=nnzb, vb=vb, positions=x, per_step=per_step, last=st,
                tuning=tuning, ls_fail=ls_fail, floor_fail=floor_fail, halvings=halvings)


# CG kernel shapes (lanes per block row, blocks in flight, workgroup size)
# by variant index, as k_solver.hip lists them
CG_SHAPES = [(4, 4, 256), (4, 4, 512), (8, 2, 256), (8, 2, 1024), (4, 2, 256), (4, 2, 1024),
             (16, 1, 1024), (8, 1, 256)]


def alternative_shapes(tuned_shape, n=2):
    """n CG kernel shapes other than the tuned one (other lanes per row and
    workgroup sizes, so the dot products are summed in other orders)."""
    out = []
    for v in (3, 6, 0, 2):
        if CG_SHAPES[v] != tuple(tuned_shape[:3]) and len(out) < n:
            out.append(v)
    return out


def fp64_spread_run(scene, steps, warmup, variant):
    """The FP64 run again with CG kernel shape `variant` (its dot products
    summed in another order), untimed: its final positions give the
    trajectory's sensitivity to rounding."""
    from mfem_amd import native
    sim = native.Sim(scene, precision="fp64")
    sim.set_cg_shape(variant, 0)
    for _ in range(warmup + steps):
        sim.step()
    x = sim.positions()
    sim.close()
    return x


def run_cpu_minres(scene, threads, caps=(4, 12)):
    """FP64 oracle, one Newton iteration of the same scene with MINRES capped
    at each of `caps` iterations: the two timings give the fixed cost (element
    update, line search) and the cost per MINRES iteration."""
    from oracle import oracle as O
    import copy
    out = []
    for cap in caps:
        sc = copy.copy(scene)
        sc.max_newton = 1
        sc.max_iters = cap
        sim = O.Sim(sc, threads=threads, solver=O.SOLVERS["minres-indefinite-block"])
        print(f"[bench] CPU oracle MINRES sample (cap {cap}) on {threads} threads ...",
              file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        sim.step()
        out.append(time.perf_counter() - t0)
    per_iter = (out[1] - out[0]) / (caps[1] - caps[0])
    fixed = out[0] - caps[0] * per_iter
    return fixed, per_iter, out


def run_minres(args, scene):
    """The KKT / MINRES solver ("minres-indefinite-block") on the same workload."""
    import copy
    scene = copy.copy(scene)
    scene.solver = "minres-indefinite-block"
    nT = scene.nT
    r = run_gpu(scene, "fp64", args.steps, args.warmup)
    value = r["newton"] * nT / r["elapsed"]
    ms_per_step = 1000.0 * r["elapsed"] / args.steps
    bytes_tet = minres_tet_bytes(nT, scene.nV, r["nfree"], r["vb"])
    avg_ms = r["spmv_ms"] / max(r["spmv_n"], 1)
    achieved = bytes_tet / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    its_per_newton = r["pcg"] / max(r["newton"], 1)
    cpu = None
    if not args.no_cpu:
        threads = cpu_threads(args.cpu_threads)
        fixed, per_iter, raw = run_cpu_minres(scene, threads)
        t_newton = fixed + its_per_newton * per_iter
        cpu = {"value": nT / t_newton, "unit": "Newton-steps*tets/s", "cores": threads,
               "kind": "port",
               "sample": f"FP64 oracle, 1 Newton iteration of the same {nT}-tet scene with MINRES "
                         f"capped at 4 and at 12 iterations ({raw[0]:.2f} s, {raw[1]:.2f} s): "
                         f"{fixed:.2f} s fixed + {per_iter:.3f} s per MINRES iteration, taken at "
                         f"the GPU run's {its_per_newton:.0f} iterations per Newton iteration"}
    out = {
        "metric": "Newton-steps*tets/s, KKT system with MINRES (minres-indefinite-block)",
        "value": value, "unit": "Newton-steps*tets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload_name(args.config, args.nx, args.ny, args.nz, nT, scene)
                   .replace("Jacobi CG", "MINRES (block-Schur preconditioner)"),
                   "tets": nT, "kkt_size": 3 * r["nfree"] + 12 * nT, "precision": "fp64"},
        "newton_iters": r["newton"], "minres_iters": r["pcg"],
        "minres_iter_ms": (r["last"]["solve_ms"] / max(r["last"]["pcg_iters"], 1)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_mr_tet (MINRES element kernel: update of the previous "
                               "iteration, KKT product rows of each tet)",
                     "bytes_per_launch": bytes_tet, "avg_launch_ms": avg_ms,
                     "launches": r["spmv_n"],
                     "timing": "HIP events around the kernel on its stream, first 32 iterations "
                               "of every solve"},
        "cpu_baseline": cpu,
        "phase_ms_untimed_step": {k: r["last"][k] for k in ("update_ms", "solve_ms", "ls_ms")},
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


def run_ic(args, scene):
    """CG with the multicolour block IC(0) preconditioner ("eigen-pcg-IC") on
    the same workload and precision."""
    import copy
    scene = copy.copy(scene)
    scene.solver = "eigen-pcg-IC"
    nT = scene.nT
    r = run_gpu(scene, "mixed", args.steps, args.warmup, profile=False)
    value = r["newton"] * nT / r["elapsed"]
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_entry(scene, cpu_threads(args.cpu_threads), args,
                                 label="IC-CG, serial triangular solves")
    out = {
        "metric": "Newton-steps*tets/s, Schur complement with IC-preconditioned CG (eigen-pcg-IC)",
        "value": value, "unit": "Newton-steps*tets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * r["elapsed"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": workload_name(args.config, args.nx, args.ny, args.nz, nT, scene)
                   .replace("Jacobi CG", "IC(0)-preconditioned CG"),
                   "tets": nT, "free_dofs": 3 * r["nfree"], "precision": "mixed"},
        "newton_iters": r["newton"], "pcg_iters": r["pcg"],
        "cg_iter_ms": r["last"]["solve_ms"] / max(r["last"]["pcg_iters"], 1),
        "cpu_baseline": cpu,
        "phase_ms_untimed_step": {k: r["last"][k] for k in ("update_ms", "solve_ms", "ls_ms")},
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


def cpu_threads(requested=0):
    """Host threads for the CPU baseline: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when the environment sets it (the share of the
    host given to one GPU)."""
    if requested:
        return requested
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        avail = min(avail, int(env))
    return max(1, avail)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_sample(scene, threads, steps=3, warmup=1, max_newton=None):
    """FP64 oracle on the same scene: `warmup` timesteps not counted, then the
    median over `steps` timesteps of Newton iterations x tets / seconds.
    Returns (median rate, list of (seconds, newton, cg) per measured step)."""
    from oracle import oracle as O
    import copy
    sc = copy.copy(scene)
    if max_newton is not None:
        sc.max_newton = max_newton
    sim = O.Sim(sc, threads=threads)
    print(f"[bench] CPU oracle sample on {threads} threads: {warmup} warm-up + {steps} timesteps",
          file=sys.stderr, flush=True)
    rows = []
    for k in range(warmup + steps):
        t0 = time.perf_counter()
        sim.step()
        dt = time.perf_counter() - t0
        st = sim.stats()
        print(f"[bench] CPU oracle timestep {k}{' (warm-up)' if k < warmup else ''}: {dt:.2f} s, "
              f"newton {st['newton_iters']}, cg {st['pcg_iters_total']}", file=sys.stderr,
              flush=True)
        if k >= warmup:
            rows.append((dt, st["newton_iters"], st["pcg_iters_total"]))
    rates = sorted(n * scene.nT / t for t, n, _ in rows)
    return rates[len(rates) // 2], rows


def cpu_baseline_entry(scene, threads, args, label="Jacobi CG"):
    rate, rows = run_cpu_sample(scene, threads, steps=args.cpu_steps)
    secs = ", ".join(f"{t:.2f} s/{n} Newton/{c} CG" for t, n, c in rows)
    return {"value": rate, "unit": "Newton-steps*tets/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "cpus_visible": os.cpu_count(),
            "sample": f"FP64 oracle ({label}) on the same {scene.nT}-tet scene: 1 warm-up "
                      f"timestep, then the median over {len(rows)} timesteps ({secs})"}


def run_batch(scs, precision, steps, warmup, graphs=True, shape=None):
    from mfem_amd import native
    b = native.Batch(scs, precision=precision)
    b.set_graphs(graphs)
    if shape is not None:
        b.set_cg_shape(*shape)
    for _ in range(warmup):
        b.step()
    work = 0
    pcg = 0
    t0 = time.perf_counter()
    for k in range(steps):
        ts = time.perf_counter()
        b.step()
        st = b.stats()
        work += int((st["newton_iters"].astype(np.int64) * b.nT).sum())
        pcg += int(st["pcg_iters"].sum())
        print(f"[bench] c4 {precision} step {k}: {(time.perf_counter() - ts) * 1e3:.1f} ms, "
              f"newton {int(st['newton_iters'].sum())} (max {int(st['newton_iters'].max())}), "
              f"cg {int(st['pcg_iters'].sum())} (max {int(st['pcg_iters'].max())})",
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    pos = [b.positions(k) for k in range(b.count)]
    shape = b.cg_shape()
    counts = b.launch_counts()
    b.close()
    return dict(elapsed=elapsed, work=work, pcg=pcg, positions=pos, shape=shape, counts=counts)


def run_c4(args):
    """Batching config: independent mixed beams in one batch; value counts
    Newton iterations x tets summed over the beams per second. Parity: the
    smallest beams are replayed alone in the FP64 oracle for the same steps."""
    from mfem_amd import scenes
    from oracle import oracle as O
    scs = scenes.config_c4(args.beams)
    nT = np.array([sc.nT for sc in scs])
    # the library's default (eager launches) is the value; the hipGraph replay
    # of the same batch with the same CG shape is reported beside it
    mixed = run_batch(scs, "mixed", args.steps, args.warmup, graphs=False)
    graph = run_batch(scs, "mixed", args.steps, args.warmup, graphs=True, shape=mixed["shape"])
    fp64 = None if args.no_fp64 else run_batch(scs, "fp64", args.steps, args.warmup, graphs=False)
    value = mixed["work"] / mixed["elapsed"]
    out = {
        "metric": "Newton-steps*tets/s at 1M tets, FP32-mixed vs FP64 GPU; PCG HBM GB/s",
        "value": value, "unit": "Newton-steps*tets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * mixed["elapsed"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"{len(scs)} independent mixed-stretch SNH jittered Kuhn beams, "
                               f"{int(nT.min())}-{int(nT.max())} tets each ({int(nT.sum())} total), "
                               f"E log-uniform [1e4,1e7], nu uniform [0.3,0.48], BDF1 dt 1/30, CG "
                               f"tol 1e-4 / 500 its, 5 Newton its, one batch",
                   "beams": len(scs), "tets": int(nT.sum()), "precision": "mixed"},
        "pcg_iters": mixed["pcg"], "cg_shape_variant_rows_per_workgroup": list(mixed["shape"]),
        "graphs": {
            "graph_value": graph["work"] / graph["elapsed"],
            "graph_ms_per_step": 1000.0 * graph["elapsed"] / args.steps,
            "bit_identical_to_eager": all(np.array_equal(a, b) for a, b in
                                          zip(graph["positions"], mixed["positions"])),
            "submissions_per_newton_iteration": {
                "update_phase": {"graphs": graph["counts"][0] / graph["counts"][2],
                                 "eager": mixed["counts"][0] / mixed["counts"][2]},
                "cg": {"graphs": graph["counts"][1] / graph["counts"][2],
                       "eager": mixed["counts"][1] / mixed["counts"][2]}}},
    }
    if fp64 is not None:
        v64 = fp64["work"] / fp64["elapsed"]
        out["fp64_gpu"] = {"value": v64, "ms_per_step": 1000.0 * fp64["elapsed"] / args.steps}
        out["mixed_over_fp64"] = value / v64
        errs = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(mixed["positions"], fp64["positions"])]
        out["mixed_vs_fp64_rel_pos_err_max"] = float(max(errs))
    if not args.no_cpu:
        # FP64 oracle on the 8 smallest beams for the same number of steps
        order = np.argsort(nT)[:8]
        t0 = time.perf_counter()
        work = 0
        errs = []
        for k in order:
            o = O.Sim(scs[k], threads=cpu_threads(args.cpu_threads))
            for _ in range(args.warmup + args.steps):
                o.step()
                work += o.stats()["newton_iters"] * scs[k].nT
            xo = o.positions()
            errs.append(float(np.linalg.norm(mixed["positions"][k] - xo) / np.linalg.norm(xo)))
        dt_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": work / dt_cpu, "unit": "Newton-steps*tets/s",
                               "cores": cpu_threads(args.cpu_threads),
                               "kind": "port",
                               "sample": f"FP64 oracle on the 8 smallest beams, "
                                         f"{args.warmup + args.steps} steps each, {dt_cpu:.2f} s"}
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        out["mixed_vs_oracle_rel_pos_err_max_8_smallest"] = float(max(errs))
    print(json.dumps(out), flush=True)


def workload_name(config, nx, ny, nz, nT, scene):
    beam = f"jittered Kuhn beam {nx}x{ny}x{nz}x6 ({nT} tets)"
    if config == "c3":
        return (f"mixed-stretch SNH {beam}, per-tet E log-uniform [1e5,1e6], nu 0.45, BDF1 dt 1/30, "
                f"Jacobi CG tol 1e-4 / 500 its, 5 Newton its")
    if config == "c2":
        return (f"plain stretch SNH {beam}, E 1e6, nu 0.45, BDF1 dt 1/30, Jacobi CG rel-res 1e-10 / "
                f"20000 its, 5 Newton its")
    return (f"mixed-stretch {beam}, ARAP / SNH split by a seeded plane "
            f"({100 * scene.extra['arap_fraction']:.1f}% ARAP), per-tet E log-uniform [1e4,1e8], "
            f"nu 0.49, BDF1 dt 1/30, Jacobi CG tol 1e-4 / 500 its, 5 Newton its")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", choices=("c3", "c2", "c5", "c4"), default="c3",
                    help="c3: the metric's 960k-tet mixed beam (default); c2: plain FP64 "
                         "24k-tet beam; c5: 1.95M-tet stiff ARAP/SNH beam; c4: 256 independent "
                         "beams (2e3-6e4 tets each) as one batch")
    ap.add_argument("--beams", type=int, default=256, help="c4: number of beams")
    ap.add_argument("--nx", type=int, default=0)
    ap.add_argument("--ny", type=int, default=0)
    ap.add_argument("--nz", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="timesteps of the CPU-baseline median (after one warm-up timestep)")
    ap.add_argument("--solver", choices=("eigen-pcg-diag", "eigen-pcg-IC",
                                         "minres-indefinite-block"),
                    default="eigen-pcg-diag",
                    help="linear solver of the Newton step: Schur complement + Jacobi CG "
                         "(default, the metric's), Schur complement + IC(0) CG, or the KKT "
                         "system with MINRES (FP64)")
    ap.add_argument("--element-kernel", type=int, choices=(0, 1), default=0,
                    help="mixed-precision run: stiffness-block products on the vector units (0, "
                         "default) or on the matrix cores (1, MFMA)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    from mfem_amd import scenes
    if args.config == "c4":
        return run_c4(args)
    dims = {"c3": (100, 40, 40), "c2": (40, 10, 10), "c5": (130, 50, 50)}[args.config]
    dims = tuple(a or d for a, d in zip((args.nx, args.ny, args.nz), dims))
    args.nx, args.ny, args.nz = dims
    make = {"c3": scenes.config_c3, "c2": scenes.config_c2, "c5": scenes.config_c5}[args.config]
    scene = make(*dims)
    nT = scene.nT
    if args.solver == "minres-indefinite-block":
        return run_minres(args, scene)
    if args.solver == "eigen-pcg-IC":
        return run_ic(args, scene)

    mixed = run_gpu(scene, "mixed", args.steps, args.warmup, element_kernel=args.element_kernel)
    fp64 = None if args.no_fp64 else run_gpu(scene, "fp64", args.steps, args.warmup)

    value = mixed["newton"] * nT / mixed["elapsed"]
    ms_per_step = 1000.0 * mixed["elapsed"] / args.steps
    bytes_spmv = cg_bytes(mixed["nfree"], mixed["nnzb"], mixed["vb"])
    avg_spmv_ms = mixed["spmv_ms"] / max(mixed["spmv_n"], 1)
    achieved = bytes_spmv / (avg_spmv_ms * 1e-3) / 1e9 if avg_spmv_ms > 0 else 0.0

    extra = {}
    valid = True
    reasons = []
    extra["mixed_ls_failures"] = sum(mixed["ls_fail"])
    extra["mixed_ls_failures_at_rounding_floor"] = sum(mixed["floor_fail"])
    extra["mixed_ls_halvings"] = sum(mixed["halvings"])
    if fp64 is not None:
        v64 = fp64["newton"] * nT / fp64["elapsed"]
        x64 = fp64["positions"]
        err = float(np.linalg.norm(mixed["positions"] - x64) / np.linalg.norm(x64))
        avg64 = fp64["spmv_ms"] / max(fp64["spmv_n"], 1)
        b64 = cg_bytes(fp64["nfree"], fp64["nnzb"], fp64["vb"])
        extra["fp64_gpu"] = {
            "value": v64, "ms_per_step": 1000.0 * fp64["elapsed"] / args.steps,
            "newton_iters": fp64["newton"], "pcg_iters": fp64["pcg"],
            "ls_failures": sum(fp64["ls_fail"]), "ls_halvings": sum(fp64["halvings"]),
            "cg_kernel_gbs": b64 / (avg64 * 1e-3) / 1e9 if avg64 > 0 else 0.0,
            "cg_kernel_ms": avg64}
        extra["mixed_over_fp64"] = value / v64 if v64 > 0 else None
        extra["mixed_vs_fp64_rel_pos_err"] = err
        # rounding floor: FP64 against FP64 with other dot-product orders
        tuned = (fp64["tuning"]["lanes_per_row"], fp64["tuning"]["blocks_in_flight"],
                 fp64["tuning"]["workgroup"]) if fp64["tuning"] else (4, 4, 256)
        # two more FP64 runs with other shapes: one pair is one sample of the
        # trajectory's rounding sensitivity; the largest of the three pairs
        # is the spread
        runs = [x64] + [fp64_spread_run(scene, args.steps, args.warmup, v)
                        for v in alternative_shapes(tuned)]
        pairs = [float(np.linalg.norm(a - b) / np.linalg.norm(x64))
                 for i, a in enumerate(runs) for b in runs[i + 1:]]
        spread = max(pairs)
        extra["fp64_vs_fp64_rounding_spread"] = spread
        extra["fp64_vs_fp64_pairs"] = pairs
        bar = max(1e-6, 3.0 * spread)
        extra["mixed_error_bar"] = bar
        # the bar is the precision config's (c3); c2 and c5 report
        if args.config == "c3" and not err <= bar:
            valid = False
            reasons.append(f"mixed positions {err:.3e} from FP64 (bar {bar:.1e}: 1e-6 or 3x the "
                           f"FP64 rounding spread {spread:.1e})")
        extra_fail = sum(mixed["ls_fail"]) - sum(mixed["floor_fail"])
        if args.config == "c3" and extra_fail > sum(fp64["ls_fail"]):
            valid = False
            reasons.append(f"mixed failed {extra_fail} line searches before reaching rounding, "
                           f"FP64 {sum(fp64['ls_fail'])}")
    extra["valid"] = valid
    if reasons:
        extra["invalid_reason"] = "; ".join(reasons)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_entry(scene, cpu_threads(args.cpu_threads), args)
        extra["gpu_over_cpu"] = value / cpu["value"] if cpu["value"] > 0 else None

    out = {
        "metric": "Newton-steps*tets/s at 1M tets, FP32-mixed vs FP64 GPU; PCG HBM GB/s",
        "value": value,
        "unit": "Newton-steps*tets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": workload_name(args.config, args.nx, args.ny, args.nz, nT, scene),
                   "tets": nT, "free_dofs": 3 * mixed["nfree"], "bsr_blocks": mixed["nnzb"],
                   "precision": "mixed"},
        "newton_iters": mixed["newton"],
        "pcg_iters": mixed["pcg"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(mixed["tuning"]["lanes_per_row"],
                                                 mixed["tuning"]["blocks_in_flight"],
                                                 mixed["tuning"]["workgroup"],
                                                 args.config == "c3" and dims == (100, 40, 40)),
                     "traffic_source": "profiles/k_cg_iter_c3_mixed_traffic.md (rocprofv3 FETCH_SIZE x 2.0, "
                                       "WRITE_SIZE x 1.0 from a copy calibration)",
                     "kernel": "k_cg_iter (one CG iteration: BSR 3x3 SpMV, x/r/z/p updates, "
                               "5 dot-product partials)",
                     "bytes_per_launch": bytes_spmv, "avg_launch_ms": avg_spmv_ms,
                     "launches": mixed["spmv_n"],
                     "timing": "per-workgroup device wall-clock stamps, span of each launch"},
        "cpu_baseline": cpu,
        "phase_ms_untimed_step": {k: mixed["last"][k] for k in ("update_ms", "solve_ms", "ls_ms")},
        "cg_autotune": mixed["tuning"],
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    if not valid:
        print(f"[bench] INVALID: {extra['invalid_reason']}", file=sys.stderr, flush=True)
        sys.exit(3)


if __name__ == "__main__":
    main()
