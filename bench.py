"""Benchmark: converged states/s of the ITP loop on one MI355X, mixed vs FP64.

Default workload (the batched configuration): 64 independent quantum-dot
instances (grids 128^2/192^2/256^2, 50-400 states, oscillator plus 40 Gaussian
impurities, B in [0, 2), periodic, 4th-order multi-product split, eps 0.05,
energy-change tests at 1e-9), all solved together by the batch API (worker
threads, one HIP stream each). A step is one complete batched solve of all 64
instances (seeded noise initial states, propagate -> orthonormalize ->
energies until every instance's tests pass), inputs resident in HBM, per-solve
setup inside the timed region as part of each instance (the reference's own
timer starts at the first step, itpsystem.cpp:416-421; here the whole batch
wall time counts).

value: converged states per second of the mixed-precision path, counting a
state only if the FP64 path converges it too and the two energies agree to
1e-8 relative. The FP64 path runs the same batch after the same W warm-up
solves, once (timed, its converged states/s in "fp64"); speedup_vs_fp64 is
the ratio. Both paths tune the overlap's launch shape (--autotune 1, the
default) in their warm-up solves. The 256x256 Dirichlet
single-instance workload (N = 500, B = 1) is reported beside it ("c2").

Prints one JSON line. Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged states/s @1e-8 rel eigenvalue err, mixed vs FP64 GPU path; %HBM"
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6


def measured_traffic(workload, path):
    """HBM bytes per sweep launch from the committed PMC summary of this
    workload (FETCH_SIZE and WRITE_SIZE runs, copy-calibrated), or None."""
    f = os.path.join(ROOT, "profiles", f"sweep_traffic_{workload}_{path}.md")
    if not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sweep_transfers(cfg):
    """Field transfers (N*M*s bytes each) per propagate call and per energies call."""
    h = cfg["halforder"]
    p_prop = 2 * h * h + 5 * h - 1  # 2t+1 sweeps of 2 transfers per member, plus h-1 accumulator reads
    if cfg["B"] != 0:
        p_en = 10   # per H: row (2) + column with add and dot (3); twice, the last without a write
        n_en = 4
    else:
        p_en = 14   # per H: row (2) + column (2) + row with V*psi and dot (3); twice
        n_en = 6
    return p_prop, h * h + 2 * h, p_en, n_en


def cpu_baseline(cfg, V, per_solve_iters, per_solve_converged, sample_steps):
    from oracle import pyoracle as O
    threads = O.num_threads()
    sysc = O.System(cfg, V)
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        sysc.step()
    dt = time.perf_counter() - t0
    s_per_iter = dt / sample_steps
    value = per_solve_converged / (per_solve_iters * s_per_iter)
    return {"value": value, "unit": "converged states/s", "cores": threads, "kind": "port",
            "sample": f"{sample_steps} ITP iteration(s) of the same workload with the FP64 C++ oracle "
                      f"({s_per_iter:.2f} s/iteration, OMP threads {threads}), scaled by the GPU solve's "
                      f"{per_solve_iters} iterations and {per_solve_converged} converged states"}


def mixed_leg(I, ctx, cfg, V, args):
    """The mixed-precision path on the same workload: timed solves, and the
    converged states whose energies agree with the FP64 path's (one FP64 solve
    of the same seeded input, run here as the reference) to 1e-8 relative."""
    ref = I.ITPSystem(ctx, cfg, V)
    ref.run()
    E64 = ref.sorted()[0]
    conv64 = ref.converged()
    del ref
    mcfg = dict(cfg, precision=I.FP32)
    for w in range(args.warmup):
        s = I.ITPSystem(ctx, mcfg, V)
        s.run()
        del s
    systems = [I.ITPSystem(ctx, mcfg, V) for _ in range(args.steps)]
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in systems:
        s.run()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    good, worst = 0, 0.0
    for s in systems:
        E = s.sorted()[0]
        c = s.converged()
        err = np.abs(E - E64) / np.abs(E64)
        ok = c & (err <= 1e-8)
        good += int(ok.sum())
        if c.any():
            worst = max(worst, float(err[c].max()))
    s0 = systems[0]
    out = {"value": round(good / wall, 3), "unit": "converged states/s", "ms_per_step": round(1e3 * wall / len(systems), 3),
           "steps_per_solve": s0.total_steps, "fp32_steps_per_solve": s0.mixed_steps,
           "converged_per_solve": s0.num_converged, "converged_within_1e-8_of_fp64": good // len(systems),
           "fp64_converged": int(conv64.sum()), "max_rel_err_vs_fp64": worst,
           "dtype": "c32 fields and sweeps, bf16x3 MFMA overlap (fp32 accumulate, fp64 sums), "
                    "split-bf16x3 MFMA rotation (fp32 accumulate), "
                    "fp64 eigensolver; fp64 final steps"}
    log(f"mixed leg: {out}")
    return out


def copy_probe(mib):
    """Three device-to-device copies of a known size: the per-dispatch FETCH_SIZE and WRITE_SIZE
    they read give the factor that converts those counters to bytes on this device."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "dispatches": 3}), flush=True)


def batch_solve(I, insts, prec, workers, autotune=False):
    """One batched solve of all instances on one path: (rc, wall s, per-instance results)."""
    b = I.Batch(0, workers=workers)
    b.set_autotune(int(autotune))
    for c, V in insts:
        b.add(dict(c, precision=prec), V)
    rc = b.run()
    res = []
    for i in range(len(insts)):
        r = b.result(i)
        E, _, f = b.sorted(i)
        r.update(E=E, conv=f, mixed=b.mixed_stats(i))
        res.append(r)
    wall = b.wall_time
    b.close()
    return rc, wall, res


def agreement(res, ref, tol=1e-8):
    """(states converged on both paths within tol of the reference's energies,
    converged on both but farther, converged here only, largest relative error
    among the states converged on both)."""
    good = far = extra = 0
    worst = 0.0
    for r, q in zip(res, ref):
        both = r["conv"] & q["conv"]
        err = np.abs(r["E"] - q["E"]) / np.abs(q["E"])
        good += int((both & (err <= tol)).sum())
        far += int((both & (err > tol)).sum())
        extra += int((r["conv"] & ~q["conv"]).sum())
        if both.any():
            worst = max(worst, float(err[both].max()))
    return good, far, extra, worst


def sweep_roofline(I, cfg, V, calls=20):
    """Roofline of the dominant kernel group (the FP32 sweeps of propagation) at
    the shape of one instance: `calls` propagations of its state set on the
    context's stream, timed with the library's HIP events around each call."""
    prec = I.FP32
    N, M = cfg["N"], cfg["sx"] * cfg["sy"]
    with I.Context(0) as ctx:
        dx = cfg["lenx"] / cfg["sx"]
        g = I.Grid(ctx, cfg["sx"], cfg["sy"], dx, cfg["boundary"])
        pot = I.Potential(g, V)
        sp = I.MultiProductSplit(pot, cfg["halforder"], cfg["eps_values"][0], cfg["B"])
        S = I.StateSet(g, N, prec)
        S.init_noise(cfg["seed"], dx)
        sp.propagate(S)
        ctx.synchronize()
        ctx.enable_timing(True)
        ctx.reset_phase_times()
        for _ in range(calls):
            sp.propagate(S)
        ctx.synchronize()
        ms, cnt = ctx.phase_times()["propagate"]
        ctx.enable_timing(False)
        h = sp.halforder
    p_prop = 2 * h * h + 5 * h - 1
    launches = h * h + 2 * h
    field = N * M * 8
    per_launch_bytes = p_prop * field / launches
    avg_ms = ms / (cnt * launches)
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    return {"kernel": "reg_sweep_kernel<float> (propagation sweeps)", "bound": "hbm",
            "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_ms": round(avg_ms, 5), "algorithmic_bytes_per_launch": round(per_launch_bytes),
            "shape": f"{cfg['sx']}x{cfg['sy']}, N={N}, h={h}, complex FP32"}


def largest_instance(insts):
    return max(insts, key=lambda x: x[0]["N"] * x[0]["sx"] * x[0]["sy"])


def c4_cpu_baseline(insts, index=0):
    """The FP64 C++ oracle's full solve of one instance (the smallest) on the host cores."""
    from oracle import pyoracle as O
    cfg, V = insts[index]
    threads = O.num_threads()
    sysc = O.System(cfg, V)
    t0 = time.perf_counter()
    sysc.run()
    dt = time.perf_counter() - t0
    conv = sysc.num_converged
    return {"value": round(conv / dt, 4), "unit": "converged states/s", "cores": threads, "kind": "port",
            "sample": f"full FP64 solve of instance {index} ({cfg['sx']}x{cfg['sy']}, N={cfg['N']}, "
                      f"{sysc.total_steps} ITP iterations, {conv} converged) with the FP64 C++ oracle, "
                      f"{dt:.2f} s on {threads} OpenMP threads"}


def c2_summary(I, W):
    """The 256x256 Dirichlet workload, one solve per path (the FP64 solve is the reference)."""
    cfg, V = W.config("c2_dirichlet256")
    out = {}
    with I.Context(0) as ctx:
        walls, res = {}, {}
        for name, prec in [("fp64", I.FP64), ("mixed", I.FP32)]:
            s = I.ITPSystem(ctx, dict(cfg, precision=prec), V)
            s.run()  # warm-up
            s.close()
            s = I.ITPSystem(ctx, dict(cfg, precision=prec), V)
            ctx.synchronize()
            t0 = time.perf_counter()
            s.run()
            ctx.synchronize()
            walls[name] = time.perf_counter() - t0
            res[name] = {"E": s.sorted()[0], "conv": s.converged(), "steps": s.total_steps,
                         "mixed": s.mixed_stats()}
            s.close()
    good, far, extra, worst = agreement([res["mixed"]], [res["fp64"]])
    v64 = int(res["fp64"]["conv"].sum()) / walls["fp64"]
    vm = good / walls["mixed"]
    out = {"workload": "c2_dirichlet256", "grid": "256x256 dirichlet", "N": cfg["N"], "B": cfg["B"],
           "halforder": cfg["halforder"],
           "fp64": {"value": round(v64, 3), "ms_per_solve": round(walls["fp64"] * 1e3, 3),
                    "steps": res["fp64"]["steps"], "converged": int(res["fp64"]["conv"].sum())},
           "mixed": {"value": round(vm, 3), "ms_per_solve": round(walls["mixed"] * 1e3, 3),
                     "steps": res["mixed"]["steps"], "fp32_steps": res["mixed"]["mixed"][0],
                     "converged_within_1e-8": good, "converged_farther": far, "converged_extra": extra,
                     "max_rel_err_vs_fp64": worst},
           "speedup_vs_fp64": round(vm / v64, 3)}
    log(f"c2: {out}")
    return out


def c4_headline(I, W, args):
    insts = W.c4_instances(64)
    states = sum(c["N"] for c, _ in insts)
    log(f"c4: {len(insts)} instances, {states} states, workers {args.batch_workers}, autotune {args.autotune}")
    tune = bool(args.autotune)
    # the FP64 path: warm-up solves as for the mixed path (kernel loading, launch-shape tuning), then
    # one timed solve, which also gives the reference energies of the mixed path
    for w in range(args.warmup):
        rc, wall, _ = batch_solve(I, insts, I.FP64, args.batch_workers, tune)
        log(f"fp64 warm-up {w}: rc={rc} {wall:.2f}s")
    rc64, wall64, res64 = batch_solve(I, insts, I.FP64, args.batch_workers, tune)
    conv64 = sum(int(r["conv"].sum()) for r in res64)
    it64 = sum(r["steps"] for r in res64)
    log(f"fp64 batch rc={rc64}: {wall64:.2f}s, {conv64} converged, {it64} ITP iterations")
    for w in range(args.warmup):
        rc, wall, _ = batch_solve(I, insts, I.FP32, args.batch_workers, tune)
        log(f"mixed warm-up {w}: rc={rc} {wall:.2f}s")
    walls, goods, fars, extras, worst, its, fp32, checks, skips, fb = [], [], [], [], 0.0, 0, 0, 0, 0, 0
    for k in range(args.steps):
        rc, wall, res = batch_solve(I, insts, I.FP32, args.batch_workers, tune)
        good, far, extra, w_ = agreement(res, res64)
        walls.append(wall)
        goods.append(good)
        fars.append(far)
        extras.append(extra)
        worst = max(worst, w_)
        its = sum(r["steps"] for r in res)
        fp32 = sum(r["mixed"][0] for r in res)
        checks = sum(r["mixed"][1] for r in res)
        fb = sum(r["mixed"][2] for r in res)
        skips = sum(r["mixed"][3] for r in res)
        failed = [i for i, r in enumerate(res) if r["status"] or r["error_flag"]]
        log(f"mixed step {k}: rc={rc} {wall:.2f}s, {good} converged within 1e-8 of fp64 ({far} farther, "
            f"{extra} converged only here), failed instances {failed}")
    total = sum(walls)
    value = sum(goods) / total
    v64 = conv64 / wall64
    roof = sweep_roofline(I, *largest_instance(insts))
    roof["traffic"] = measured_traffic("c4_largest", "fp32")
    log(f"roofline: {roof}")
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "converged states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c32 fields; bf16x3 MFMA Gram and rotation with fp32 accumulation; f64 Cholesky, "
                 "eigensolver, checks and final steps",
        "data": "synthetic (64 seeded instances: grids 128^2/192^2/256^2, 50-400 states, B in [0,2), "
                "oscillator plus 40 Gaussian impurities, seeded noise initial states)",
        "config": {"workload": "c4_batch64", "instances": len(insts), "states": states,
                   "workers": args.batch_workers, "path": "mixed", "autotune": tune,
                   "converged_within_1e-8_of_fp64": goods[-1], "converged_farther_than_1e-8": fars[-1],
                   "converged_not_by_fp64": extras[-1], "max_rel_err_vs_fp64": worst,
                   "itp_iterations": its, "fp32_iterations": fp32, "fp64_checks": checks,
                   "skipped_orthonormalizations": skips, "cholesky_fallbacks": fb,
                   "itp_iterations_per_s": round(its / (total / args.steps), 3)},
        "fp64": {"value": round(v64, 3), "unit": "converged states/s", "ms_per_step": round(wall64 * 1e3, 3),
                 "converged": conv64, "itp_iterations": it64, "dtype": "f64"},
        "speedup_vs_fp64": round(value / v64, 3),
        "roofline": roof,
    }
    if not args.no_c2:
        line["c2"] = c2_summary(I, W)
    if not args.no_cpu_baseline:
        log("cpu baseline sample ...")
        line["cpu_baseline"] = c4_cpu_baseline(insts)
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4_batch64")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--cpu-sample-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--iterations", type=int, default=0,
                    help="profiling mode: run this many ITP iterations of one solve and print the phase times "
                         "instead of the benchmark line")
    ap.add_argument("--roofline-only", action="store_true",
                    help="profiling mode: only the sweep roofline measurement of the batched configuration "
                         "(20 propagations at its largest instance's shape) and exit")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="profiling mode: run 3 plain device copies of this many MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known byte count) and exit")
    ap.add_argument("--batch-workers", type=int, default=8,
                    help="batched configuration (--workload c4_batch64): worker threads, one stream each")
    ap.add_argument("--no-c2", action="store_true", help="batched configuration: skip the c2 summary")
    ap.add_argument("--autotune", type=int, default=1, choices=[0, 1],
                    help="batched configuration: launch-shape tuning of the overlap (done in the warm-up solves)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the single-GPU throughput is the reported value)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return

    import itp2d_amd as I
    from itp2d_amd import workloads as W

    if args.workload == "c4_batch64" and args.roofline_only:
        insts = W.c4_instances(64)
        print(json.dumps(sweep_roofline(I, *largest_instance(insts))), flush=True)
        return
    if args.workload == "c4_batch64":
        c4_headline(I, W, args)
        return
    cfg, V = W.config(args.workload)
    cfg = dict(cfg, precision=I.FP32 if args.precision == "fp32" else I.FP64)
    N, M = cfg["N"], cfg["sx"] * cfg["sy"]
    esz = 8 if args.precision == "fp32" else 16
    ctx = I.Context(0)
    log(f"workload {args.workload}: {cfg['sx']}x{cfg['sy']}, N={N}, h={cfg['halforder']}, B={cfg['B']}, "
        f"{args.precision}; CUs={ctx.num_cus}")

    if args.iterations:
        t0 = time.perf_counter()
        s = I.ITPSystem(ctx, cfg, V)
        ctx.synchronize()
        log(f"setup {time.perf_counter() - t0:.2f}s")
        ctx.enable_timing(True)
        ctx.reset_phase_times()
        walls = []
        for k in range(args.iterations):
            t0 = time.perf_counter()
            s.step()
            ctx.synchronize()
            walls.append(time.perf_counter() - t0)
            log(f"iteration {k}: {walls[-1] * 1e3:.1f} ms")
        print(json.dumps({"workload": args.workload, "path": args.precision, "iterations": args.iterations,
                          "iteration_ms": [round(w * 1e3, 2) for w in walls],
                          "phase_ms": ctx.phase_times(), "eigensolver": ctx.eig_stats(),
                          "refinement": ctx.oa_stats()}), flush=True)
        ctx.close()
        return

    mixed = None
    if args.precision == "fp64" and not args.no_mixed:
        mixed = mixed_leg(I, ctx, cfg, V, args)

    for w in range(args.warmup):
        s = I.ITPSystem(ctx, cfg, V)
        rc = s.run()
        log(f"warmup solve {w}: rc={rc} steps={s.total_steps} converged={s.num_converged} "
            f"error={s.error_flag} {s.total_time:.2f}s")
        del s

    systems = [I.ITPSystem(ctx, cfg, V) for _ in range(args.steps)]
    ctx.enable_timing(True)
    ctx.reset_phase_times()
    eig0, oa0 = ctx.eig_stats(), ctx.oa_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    for k, s in enumerate(systems):
        s.run()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    phases = ctx.phase_times()
    ctx.enable_timing(False)
    eig1, oa1 = ctx.eig_stats(), ctx.oa_stats()

    iters = [s.total_steps for s in systems]
    conv = [s.num_converged for s in systems]
    errs = [s.error_flag for s in systems]
    for k, s in enumerate(systems):
        log(f"timed solve {k}: steps={iters[k]} converged={conv[k]} error={errs[k]} {s.total_time:.3f}s")
    value = sum(conv) / wall
    E0 = systems[0].sorted()[0]

    # roofline of the dominant kernel group, from the HIP events of the timed region
    p_prop, n_prop, p_en, n_en = sweep_transfers(cfg)
    t_prop, c_prop = phases["propagate"]
    t_en, c_en = phases["energies"]
    t_ov, c_ov = phases["overlap"]
    t_eig, c_eig = phases["eigensolve"]
    t_lc, c_lc = phases["lincomb"]
    field = N * M * esz
    sweep_bytes = c_prop * p_prop * field + c_en * p_en * field
    sweep_ms = t_prop + t_en
    sweep_launches = c_prop * n_prop + c_en * n_en
    gram_flops = 8.0 * M * N * (N + 1) / 2 * c_ov
    rot_flops = 8.0 * M * N * N * c_lc
    groups = {
        "sweep_kernel": dict(bound="hbm", ms=sweep_ms, achieved=sweep_bytes / (sweep_ms * 1e-3) / 1e9,
                             peak=HBM_PEAK_GBS, unit="GB/s", launches=sweep_launches,
                             per_launch_bytes=sweep_bytes / max(sweep_launches, 1)),
        "gram_kernel": dict(bound="mfma", ms=t_ov, achieved=gram_flops / (t_ov * 1e-3) / 1e12 if t_ov else 0,
                            peak=FP64_MFMA_PEAK_TFLOPS, unit="TFLOP/s", launches=c_ov),
        "rotate_kernel": dict(bound="mfma", ms=t_lc, achieved=rot_flops / (t_lc * 1e-3) / 1e12 if t_lc else 0,
                              peak=FP64_MFMA_PEAK_TFLOPS, unit="TFLOP/s", launches=c_lc),
    }
    dom = max(groups, key=lambda k: groups[k]["ms"])
    g = groups[dom]
    traffic = measured_traffic(args.workload, args.precision) if dom == "sweep_kernel" else None
    roofline = {"kernel": dom, "bound": g["bound"], "achieved": round(g["achieved"], 3), "peak": g["peak"],
                "unit": g["unit"], "frac": round(g["achieved"] / g["peak"], 4), "traffic": traffic,
                "avg_launch_ms": round(g["ms"] / max(g["launches"], 1), 5)}
    if dom == "sweep_kernel":
        roofline["algorithmic_bytes_per_launch"] = round(g["per_launch_bytes"])
    others = {k: {"achieved": round(v["achieved"], 3), "unit": v["unit"], "frac": round(v["achieved"] / v["peak"], 4),
                  "ms": round(v["ms"], 3)} for k, v in groups.items() if k != dom}

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "converged states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "c32-fields/f64-dense",
        "data": "synthetic (seeded uniform-noise initial states, seeded impurity potential)",
        "config": {
            "workload": args.workload,
            "grid": f"{cfg['sx']}x{cfg['sy']}", "boundary": "dirichlet" if cfg["boundary"] else "periodic",
            "N": N, "n_wanted": cfg["n_wanted"], "halforder": cfg["halforder"], "B": cfg["B"],
            "eps": cfg["eps_values"], "tests": [list(cfg["timestep_test"]), list(cfg["final_test"])],
            "path": args.precision,
            "itp_iterations_per_solve": iters[0], "converged_per_solve": conv[0],
            "itp_iterations_per_s": round(sum(iters) / wall, 3),
            "E_lowest": [round(x, 10) for x in E0[:4].tolist()],
            "phase_ms": {k: round(v[0], 3) for k, v in phases.items()},
            "phase_calls": {k: v[1] for k, v in phases.items()},
            "eigensolver": {"calls": eig1[0] - eig0[0], "jacobi_sweeps": eig1[1] - eig0[1],
                            "refined": oa1[0] - oa0[0], "fallbacks": oa1[1] - oa0[1]},
        },
        "roofline": roofline,
        "roofline_other": others,
    }
    if mixed is not None:
        mixed["speedup_vs_fp64"] = round(mixed["value"] / value, 3) if value else None
        line["mixed"] = mixed
    if not args.no_cpu_baseline:
        log("cpu baseline sample ...")
        line["cpu_baseline"] = cpu_baseline(W.config(args.workload)[0], V, iters[0], conv[0], args.cpu_sample_steps)
    ctx.close()  # device objects freed before the interpreter exits
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
