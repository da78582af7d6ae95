"""Benchmark of the PFC spectral stepper on one MI355X.

Workload (the headline metric "Mpoint-steps/s batched FP32 vs plain FP64"):
  * batched Misfit sweep: 64 instances = u0 {0.02,0.04,0.06,0.08} x con0
    {0.1,0.2,0.3,0.4} x 4 seeds, grids 16 each of 128^2, 256^2, 384^2, 512^2,
    FP32 (FP64 k=0 modes and reductions), hipGraph-captured steps; one
    "step" advances all 64 instances by one time step.  -> `value`
  * plain FP64 path: one Misfit 512^2 instance, eager launches.  -> fp64_plain
Initial fields are the reference initialisers (synthetic, seeded).

The dominant kernel (column pass of the 512^2 class) is timed with HIP
events on its own stream inside the timed region; its algorithmic bytes per
launch over that duration give roofline.achieved.  The CPU baseline times the
oracle (the FP64 CPU restatement, single thread) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "Mpoint-steps/s batched FP32 vs plain FP64 @ stated tol; p50/p99 sweep-job latency"
HBM_PEAK_GBS = 8000.0


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def sweep_instances():
    """(grid, u0, con0, seed) for the 64-instance sweep; grid by seed index."""
    grids = [128, 256, 384, 512]
    out = []
    for u0 in (0.02, 0.04, 0.06, 0.08):
        for con0 in (0.1, 0.2, 0.3, 0.4):
            for s in range(4):
                out.append((grids[s], u0, con0, 20200604 + s))
    return out


def build_sweep(hip, graph_chunk):
    """One batched context per grid class (16 instances each)."""
    sims = {}
    for g in (128, 256, 384, 512):
        members = [x for x in sweep_instances() if x[0] == g]
        p = hip.default_params(model=hip.MISFIT, L=g, M=g, precision=hip.FP32)
        sim = hip.Sim(p, batch=len(members))
        for i, (_, u0, con0, seed) in enumerate(members):
            q = hip.default_params(model=hip.MISFIT, L=g, M=g, precision=hip.FP32, u0=u0, con0=con0, seed=seed)
            phi, con = hip.init_fields(q)
            sim.set_fields(phi, con, inst=i)
        if graph_chunk:
            sim.set_graph(graph_chunk)
        sims[g] = sim
    return sims


def cpu_baseline(budget_s=15.0):
    """Oracle (FP64 CPU restatement, one thread, all 14 reference FFTs per
    step) on Misfit 256^2 with the GUI default flags."""
    orc = _load("pfc_oracle_bench", os.path.join(ROOT, "oracle", "oracle.py"))
    m = orc.model(orc.MISFIT, 256, 256, 1)
    sim = orc.Sim(m)
    sim.run(2, faithful=1)  # warm caches
    steps = 0
    t0 = time.perf_counter()
    while True:
        sim.run(5, faithful=1)
        steps += 5
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 200:
            break
    return {"value": 256 * 256 * steps / el / 1e6, "unit": "Mpoint-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle Misfit 256x256, {steps} steps (faithful: 14 c2c FFTs/step), {el:.1f} s, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--graph-chunk", type=int, default=25)
    ap.add_argument("--fp64-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    hip = _load("pfc_hip", os.path.join(ROOT, "pfc-exp-gui_amd", "pfc_hip.py"))
    hip.load()

    # ---- batched FP32 sweep ------------------------------------------------
    sims = build_sweep(hip, args.graph_chunk)
    pts = sum(s.batch * s.P for s in sims.values())
    for s in sims.values():
        s.step(args.warmup)
    for s in sims.values():
        s.sync()
    big = sims[512]
    big.time_kernel(hip.KID_COL, True)
    t0 = time.perf_counter()
    done = 0
    chunk = max(1, args.graph_chunk)
    while done < args.steps:
        n = min(chunk, args.steps - done)
        for s in sims.values():
            s.step(n)
        done += n
    for s in sims.values():
        s.sync()
    el = time.perf_counter() - t0
    col_ms, col_n = big.kernel_time()
    big.time_kernel(hip.KID_COL, False)
    col_bytes = big.kernel_bytes(hip.KID_COL)
    value = pts * args.steps / el / 1e6
    achieved = col_bytes / (col_ms * 1e-3) / 1e9 if col_ms > 0 else 0.0

    # ---- plain FP64 path (one instance, eager) -------------------------------
    p = hip.default_params(model=hip.MISFIT, L=512, M=512, precision=hip.FP64)
    phi, con = hip.init_fields(p)
    s64 = hip.Sim(p)
    s64.set_fields(phi, con)
    s64.step(10)
    s64.sync()
    t1 = time.perf_counter()
    s64.step(args.fp64_steps)
    s64.sync()
    el64 = time.perf_counter() - t1
    fp64 = 512 * 512 * args.fp64_steps / el64 / 1e6

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: reference initialisers (INITIAL_dis), seeded",
        "config": {
            "workload": "Misfit sweep, 64 instances (4 u0 x 4 con0 x 4 seeds), 16 each of 128^2/256^2/384^2/512^2, "
                        "FP32 + FP64 k=0/reductions, hipGraph chunks of %d steps" % args.graph_chunk,
            "points_per_step": pts,
            "fp64_plain_workload": "Misfit 512^2, one instance, FP64 (reference float narrowing), eager",
        },
        "fp64_plain": {"value": round(fp64, 1), "unit": "Mpoint-steps/s", "ms_per_step": round(el64 / args.fp64_steps * 1e3, 4)},
        "fp32_batched_over_fp64_plain": round(value / fp64, 2),
        "roofline": {
            "bound": "hbm",
            "kernel": "column pass (k_col<float,512>), 16 instances of 512^2",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "launches_timed": col_n,
            "mean_launch_ms": round(col_ms, 5),
            "bytes_per_launch": col_bytes,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
