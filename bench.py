"""Benchmark: converged states/s of the ITP loop on one MI355X.

A step is one complete ITP solve of the workload (seeded initial states,
propagate -> orthonormalize -> energies until the convergence tests pass),
with all inputs resident in HBM and the per-solve setup (tables, allocations,
uploads) done before the timed region, as the reference starts its timer at
the first step (itpsystem.cpp:416-421). value = converged states of all timed
solves / wall time of the timed solves.

Default workload: the 256x256 Dirichlet grid, N = 500, oscillator + 40
Gaussian impurities, B = 1, 4th-order multi-product split, eps = 0.05, on the
plain FP64 GPU path (value). The same workload also runs on the mixed-precision
path ("mixed" in the output line): its converged states count only if their
energies agree with the FP64 path's to 1e-8 relative.

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


def batch_bench(I, W, args):
    """The batched configuration: 64 independent instances run by the batch
    API (worker threads with one stream each), full solves; value = converged
    states of all instances / wall time of the batch run. The same instances
    run by one worker give the no-overlap reference."""
    insts = W.c4_instances(64)
    prec = I.FP32 if args.precision == "fp32" else I.FP64
    insts = [(dict(c, precision=prec), V) for c, V in insts]
    log(f"batch: {len(insts)} instances, {sum(c['N'] for c, _ in insts)} states, "
        f"workers {args.batch_workers}, {args.precision}")
    warm = I.Batch(0, workers=args.batch_workers)
    for c, V in insts[: args.batch_workers]:
        warm.add(c, V)
    warm.run()
    log(f"warmup batch {warm.wall_time:.2f}s")
    del warm

    def one(workers):
        b = I.Batch(0, workers=workers)
        for c, V in insts:
            b.add(c, V)
        rc = b.run()
        res = [b.result(i) for i in range(len(insts))]
        return rc, b.wall_time, res

    rc, wall, res = one(args.batch_workers)
    conv = sum(r["converged"] for r in res)
    iters = sum(r["steps"] for r in res)
    bad = [i for i, r in enumerate(res) if r["status"] or r["error_flag"]]
    log(f"batch run rc={rc}: {wall:.2f}s, {conv} converged, {iters} iterations, failed instances {bad}")
    seq = None
    if not args.no_batch_serial:
        _, wall1, res1 = one(1)
        seq = {"workers": 1, "wall_s": round(wall1, 3),
               "value": round(sum(r["converged"] for r in res1) / wall1, 3)}
        log(f"one worker: {wall1:.2f}s")
    line = {
        "metric": METRIC,
        "value": round(conv / wall, 3),
        "unit": "converged states/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "c32-fields/f64-dense",
        "data": "synthetic (64 seeded instances: grids 128^2/192^2/256^2, 50-400 states, B in [0,2), "
                "oscillator plus impurity potentials, seeded noise initial states)",
        "config": {"workload": "c4_batch64", "instances": len(insts), "states": sum(c["N"] for c, _ in insts),
                   "path": args.precision, "workers": args.batch_workers, "converged": conv,
                   "itp_iterations": iters, "itp_iterations_per_s": round(iters / wall, 3),
                   "failed_instances": bad},
        "serial": seq,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2_dirichlet256")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--cpu-sample-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--iterations", type=int, default=0,
                    help="profiling mode: run this many ITP iterations of one solve and print the phase times "
                         "instead of the benchmark line")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="profiling mode: run 3 plain device copies of this many MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known byte count) and exit")
    ap.add_argument("--batch-workers", type=int, default=4,
                    help="batched configuration (--workload c4_batch64): worker threads, one stream each")
    ap.add_argument("--no-batch-serial", action="store_true",
                    help="batched configuration: skip the one-worker comparison run")
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

    if args.workload == "c4_batch64":
        batch_bench(I, W, args)
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
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
