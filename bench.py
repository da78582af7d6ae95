"""Benchmark of the PFC spectral stepper on one MI355X.

Workload (the headline metric "Mpoint-steps/s batched FP32 vs plain FP64
@ stated tol; p50/p99 sweep-job latency"):
  * batched Misfit sweep: 64 instances = u0 {0.02,0.04,0.06,0.08} x con0
    {0.1,0.2,0.3,0.4} x 4 seeds, grids 16 each of 128^2, 256^2, 384^2, 512^2,
    FP32 (FP64 k=0 modes and reductions), hipGraph-captured steps; one
    "step" advances all 64 instances by one time step.  -> `value`
  * plain FP64 path: the sweep's 64 instances one at a time (one-instance
    FP64 contexts with the reference narrowing, eager launches).  -> fp64_plain
  * accuracy: Misfit 512^2, 256^2 and 128^2, CVD 128x384, 256^2 and 128^2, 10^4
    steps in FP32 and in the FP64 drop-in from the same initial fields, beside
    an FP64 ensemble (no narrowing; fields one FP32 round-off away, three
    seeds) whose spread is the envelope of the FP32 bar; field L-inf, energy
    and Phimax atom count every 1000 steps (CVD 128x384 also from step 10).
    -> accuracy  (The FP64 drop-in against the oracle's own trajectory over
    the same horizon is tests/oracle_curves.py: the oracle is the checker of
    the test suite, and bench.py runs it only as the CPU baseline.)
  * request loop: a 20 s stream of mixed Misfit/CVD jobs (grids 128^2,
    256^2, 512^2, 1024^2, 128x128x64; 500..20000 steps) offering 0.8 of the
    server's measured capacity, Poisson arrivals, through the
    continuous-batching server in FP32, final files written per job; the first
    jobs of the same stream also run one at a time on the plain FP64 path.
    -> serving (p50/p99 job latency, aggregate Mpoint-steps/s)
  * CVD 1024^2 in FP32, 2x10^4 steps with the energy traced every step and a
    checkpoint every 500 steps, then the elastic post-process of every
    checkpoint, all on the GPU through the product CLI.  -> cvd_checkpoints
Initial fields are the reference initialisers (synthetic, seeded).

The dominant kernel (column pass of the 512^2 class) is timed with HIP
events on its own stream inside the timed region (hipExtLaunchKernel
start/stop events: the kernel's own start and end); its algorithmic bytes per
launch over that duration give roofline.achieved.  The CPU baseline times the
oracle (the FP64 CPU restatement, single thread) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --pmc-drive   (a few sweep steps and two 256 MiB copy
                                      kernels, for hardware-counter runs)
"""
import argparse
import importlib.util
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "Mpoint-steps/s batched FP32 vs plain FP64 @ stated tol; p50/p99 sweep-job latency"
WARM_S = 0.25  # seconds of group-graph launches before the W warm-up steps
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


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


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
            "sample": f"oracle Misfit 256x256, {steps} steps (faithful: 14 c2c FFTs/step), {el:.1f} s, 1 thread",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "fft_backend": "the oracle's own FP64 mixed-radix Stockham FFT (radix 4/2/3, cached twiddles); "
                           "FFTW3, the reference's FFT library, is not installed on this host"}


def cpu_baseline_concurrent(nproc, steps=150):
    """The reference runs one single-threaded job at a time; the best case
    for a CPU host is one such job per core.  Runs `nproc` oracle processes
    (the same 256^2 job) at once and sums their rates over the timed steps."""
    code = ("import importlib.util, sys, time\n"
            "spec = importlib.util.spec_from_file_location('o', sys.argv[1])\n"
            "o = importlib.util.module_from_spec(spec); spec.loader.exec_module(o)\n"
            "s = o.Sim(o.model(o.MISFIT, 256, 256, 1)); s.run(2, faithful=1)\n"
            "t = time.perf_counter(); s.run(%d, faithful=1); print(time.perf_counter() - t)\n" % steps)
    path = os.path.join(ROOT, "oracle", "oracle.py")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    t0 = time.perf_counter()
    procs = [subprocess.Popen([sys.executable, "-c", code, path], stdout=subprocess.PIPE, env=env)
             for _ in range(nproc)]
    secs = [float(p.communicate()[0].decode().strip().splitlines()[-1]) for p in procs]
    wall = time.perf_counter() - t0
    rates = [256 * 256 * steps / t / 1e6 for t in secs]
    return {"procs": nproc, "value": sum(rates), "unit": "Mpoint-steps/s", "per_proc": sum(rates) / nproc,
            "sample": f"{nproc} oracle processes at once, Misfit 256x256, {steps} steps each, "
                      f"{min(secs):.1f}-{max(secs):.1f} s timed, {wall:.1f} s wall"}


EPS32 = 2.0 ** -24  # FP32 unit round-off


def _acc_params(hip, model, L, M, prec, narrow=1):
    kw = dict(grainx=16, grainy=16) if model == hip.MISFIT else dict(u0=0.05, con0=0.25)
    if model == hip.MISFIT and L < 256:
        kw = dict(grainx=max(1, L // 16), grainy=max(1, M // 16))
    p = hip.default_params(model=model, L=L, M=M, precision=prec, **kw)
    p.narrow = narrow
    return p


ACC_MEMBERS = ("exact", "pert", "pert2", "pert3")  # the FP64 envelope of the bar
ACC_RUNS = ACC_MEMBERS + ("pert64",)  # also recorded: the drop-in from fields 64 FP32 round-offs away


def _acc_sims(hip, model, L, M):
    """The runs of one accuracy case, from the same reference initial fields:
    'ref' the FP64 drop-in (reference narrowing), 'exact' FP64 without the
    narrowing, 'pert'/'pert2'/'pert3' the drop-in from fields changed by a
    relative +-2^-24 (one FP32 round-off per point; three seeds), 'pert64' by
    +-64 * 2^-24, 'fp32' the FP32 path."""
    sims = {}
    seeds = {"pert": (11, 1.0), "pert2": (12, 1.0), "pert3": (13, 1.0), "pert64": (14, 64.0)}
    for tag, prec, narrow in (("ref", hip.FP64, 1), ("exact", hip.FP64, 0), ("pert", hip.FP64, 1),
                              ("pert2", hip.FP64, 1), ("pert3", hip.FP64, 1), ("pert64", hip.FP64, 1),
                              ("fp32", hip.FP32, 1)):
        p = _acc_params(hip, model, L, M, prec, narrow)
        phi, con = hip.init_fields(p)
        if tag in seeds:
            rng = np.random.default_rng(seeds[tag][0])
            amp = EPS32 * seeds[tag][1]
            phi = phi * (1.0 + amp * rng.uniform(-1.0, 1.0, phi.size))
            con = con * (1.0 + amp * rng.uniform(-1.0, 1.0, con.size))
        s = hip.Sim(p)
        s.set_fields(phi, con)
        s.set_graph(100)
        sims[tag] = s
    return sims


# The FP32 bar.  The FP64 drop-in already rounds every spectral coefficient
# to float once per step (the reference's narrowing, misfit.cpp:1002-1010);
# the FP32 path rounds each value about 2 log2(P) times per step in its
# transforms (28-40 for the grids here).  So against the drop-in, at every
# checkpoint, each metric of the FP32 run may differ by at most ACC_K times
# the FP64 envelope -- the largest difference of the FP64 ensemble (the
# drop-in without the narrowing, and from fields one FP32 round-off away,
# three seeds) -- or by its floor.  Once the envelope's field distance
# reaches ACC_SATURATED the trajectory has become chaotic (two FP64
# implementations differ at O(1)); there the energy and atom count are
# compared with ACC_SPREAD times the largest FP64-ensemble difference of the
# case's saturated checkpoints so far (the ensemble's own spread).
ACC_K = 64.0
ACC_SATURATED = 1e-2
ACC_SPREAD = 4.0
ACC_FLOOR = {"linf": 1e-6, "energy_rel": 1e-6, "atoms_rel": 2e-3}
# Floors of the saturated regime: relative energy 2e-4 and atom count 1%
# (the FP32 path's larger per-step round-off seeds nucleation in a grid's
# unseeded rows earlier: Misfit 512^2 differs by 1.6e-4 in energy at 10^4
# steps, with the same atoms to 0.1%).
ACC_FLOOR_SAT = {"energy_rel": 2e-4, "atoms_rel": 1e-2}


def accuracy_case(hip, model, L, M, marks, keep_ref=False):
    """One accuracy case: FP32 and the FP64 ensemble against the FP64
    drop-in from the same initial fields, at each step count in marks."""
    sims = _acc_sims(hip, model, L, M)
    trace, refs, t = [], {}, 0
    spread = {"energy": 0.0, "atoms": 0.0}
    for mk in marks:
        for sm in sims.values():
            sm.step(mk - t)
        t = mk
        f = {k: sm.get_fields() for k, sm in sims.items()}
        e = {k: sm.energy()[0] for k, sm in sims.items()}
        if model == hip.MISFIT:
            atoms = {k: len(sm.phimax_misfit()[0]) for k, sm in sims.items()}
        else:
            atoms = {k: len(sm.phimax_cvd()[0]) for k, sm in sims.items()}
        row = {"step": t, "atoms_ref": int(atoms["ref"])}
        d = {}
        for tag in ("fp32",) + ACC_RUNS:
            lin = float(max(np.abs(f[tag][0] - f["ref"][0]).max(), np.abs(f[tag][1] - f["ref"][1]).max()))
            en = float(abs(e[tag][0] - e["ref"][0]) / max(abs(e["ref"][0]), 1e-30))
            at = abs(atoms[tag] - atoms["ref"]) / max(atoms["ref"], 1)
            d[tag] = (lin, en, at)
            row[f"linf_{tag}"], row[f"energy_rel_{tag}"], row[f"atoms_{tag}"] = lin, en, int(atoms[tag])
        env = [max(d[m][k] for m in ACC_MEMBERS) for k in range(3)]
        row["envelope"] = {"linf": env[0], "energy_rel": env[1], "atoms_rel": env[2]}
        row["saturated"] = bool(env[0] >= ACC_SATURATED)
        row["l2_fp32"] = float(np.sqrt(np.mean((f["fp32"][0] - f["ref"][0]) ** 2)))
        row["mean_phi_fp32_minus_ref"] = float(e["fp32"][1] - e["ref"][1])
        bar = {"linf": d["fp32"][0] <= max(ACC_K * env[0], ACC_FLOOR["linf"]),
               "mass": row["mean_phi_fp32_minus_ref"] == 0.0}
        if row["saturated"]:
            spread["energy"] = max(spread["energy"], env[1])
            spread["atoms"] = max(spread["atoms"], env[2])
            bar["energy"] = d["fp32"][1] <= max(ACC_SPREAD * spread["energy"], ACC_FLOOR_SAT["energy_rel"])
            bar["atoms"] = d["fp32"][2] <= max(ACC_SPREAD * spread["atoms"], ACC_FLOOR_SAT["atoms_rel"])
        else:
            bar["energy"] = d["fp32"][1] <= max(ACC_K * env[1], ACC_FLOOR["energy_rel"])
            bar["atoms"] = d["fp32"][2] <= max(ACC_K * env[2], ACC_FLOOR["atoms_rel"])
        row["bar"] = {k: bool(v) for k, v in bar.items()}
        row["bar_ok"] = all(row["bar"].values())
        if model == hip.MISFIT:
            row.update(_atom_match(sims["ref"].phimax_misfit(), sims["fp32"].phimax_misfit(), L * 0.125))
        if keep_ref:
            refs[t] = f["ref"]
        trace.append(row)
    for sm in sims.values():
        sm.close()
    return trace, refs


ACC_CASES = [(0, 512, 512), (0, 256, 256), (0, 128, 128), (1, 128, 384), (1, 256, 256), (1, 128, 128)]
EARLY_CVD = (1, 128, 384)
EARLY_MARKS = [10, 20, 50, 100, 200, 500, 1000]


def accuracy_study(hip, steps, every):
    """FP32 against the FP64 drop-in (itself pinned to the oracle by the
    test suite and tests/oracle_curves.py) from the same initial fields,
    beside the FP64 envelope, every `every` steps; and the first 1000 steps
    of CVD 128x384 at finer checkpoints."""
    marks = list(range(every, steps + 1, every))
    res = {"workload": "FP32 path and FP64 envelopes vs the FP64 drop-in (reference narrowing), every %d steps "
                       "to %d, from the reference initial fields; Misfit 512^2, 256^2 and 128^2, CVD 128x384, "
                       "256^2, 128^2; CVD 128x384 early" % (every, steps),
           "bar": {"K": ACC_K, "floor": ACC_FLOOR, "saturated": ACC_SATURATED, "spread": ACC_SPREAD,
                   "floor_saturated": ACC_FLOOR_SAT,
                   "rule": "per metric: |fp32 - ref| <= max(K * envelope, floor), envelope = the largest "
                           "|member - ref| of the FP64 ensemble (exact, pert x3); metrics: field L-inf, energy "
                           "(relative), Phimax atom count (relative); once the envelope's L-inf >= saturated, "
                           "energy and atoms within spread x the ensemble's largest difference so far; mean phi "
                           "exactly equal"},
           "cases": {}}
    for c in ACC_CASES:
        model, L, M = c
        name = f"{'misfit' if model == 0 else 'cvd'}_{L}x{M}"
        trace, _ = accuracy_case(hip, model, L, M, marks)
        res["cases"][name] = {"trace": trace, "bar_ok": all(r["bar_ok"] for r in trace)}
        print(f"accuracy {name}: final fp32 {trace[-1]['linf_fp32']:.2e} envelope "
              f"{trace[-1]['envelope']['linf']:.2e} bar_ok={res['cases'][name]['bar_ok']}", file=sys.stderr,
              flush=True)
    early, _ = accuracy_case(hip, *EARLY_CVD, EARLY_MARKS)
    res["cvd_128x384_early"] = {"trace": early, "bar_ok": all(r["bar_ok"] for r in early)}
    res["bar_ok"] = all(c["bar_ok"] for c in res["cases"].values()) and res["cvd_128x384_early"]["bar_ok"]
    return res


def _atom_match(a, b, box):
    """Phimax atoms matched by nearest neighbour (periodic box): count
    difference, largest matched position and phi differences."""
    xa, ya, pa = a[0], a[1], a[3]
    xb, yb, pb = b[0], b[1], b[3]
    out = {"atoms_fp64": int(len(xa)), "atoms_fp32": int(len(xb))}
    if len(xa) == 0 or len(xb) == 0:
        return out
    best = np.full(len(xa), np.inf)
    idx = np.zeros(len(xa), dtype=np.int64)
    for i0 in range(0, len(xb), 2048):
        dx = xa[:, None] - xb[None, i0:i0 + 2048]
        dy = ya[:, None] - yb[None, i0:i0 + 2048]
        dx -= box * np.round(dx / box)
        dy -= box * np.round(dy / box)
        d = dx * dx + dy * dy
        j = d.argmin(axis=1)
        dj = d[np.arange(len(xa)), j]
        upd = dj < best
        best[upd] = dj[upd]
        idx[upd] = j[upd] + i0
    matched = np.sqrt(best) < 0.5  # half a lattice spacing
    out["atoms_matched"] = int(matched.sum())
    out["max_pos_err"] = float(np.sqrt(best[matched]).max()) if matched.any() else None
    out["max_phi_err"] = float(np.abs(pa[matched] - pb[idx[matched]]).max()) if matched.any() else None
    return out


SERVE_GRIDS = [(128, 128, 1), (256, 256, 1), (512, 512, 1), (1024, 1024, 1), (128, 128, 64)]


def serve_jobs(hip, rng, n, prec, step_scale=1.0):
    jobs = []
    for _ in range(n):
        model = hip.MISFIT if rng.random() < 0.5 else hip.CVD
        L, M, N = SERVE_GRIDS[rng.integers(len(SERVE_GRIDS))]
        steps = max(1, int(rng.integers(500, 20001) * step_scale))
        kw = dict(u0=float(rng.uniform(0.01, 0.1)), con0=float(rng.uniform(0.1, 0.4)),
                  seed=int(rng.integers(0, 2 ** 32)))
        if model == hip.MISFIT:
            kw.update(grainx=max(1, L // 16), grainy=max(1, M // 16))
        p = hip.default_params(model=model, L=L, M=M, N=N, precision=prec, **kw)
        jobs.append((p, steps))
    return jobs


def _pct(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def serve_leg(hip, seconds, seed=20240611):
    """Poisson arrivals through the continuous-batching server (FP32), then the
    first jobs of the same stream one at a time on the plain FP64 path."""
    tmp = tempfile.mkdtemp(prefix="pfc_serve_")
    rng = np.random.default_rng(seed)
    try:
        # Capacity of the server on this mix: a closed batch of 16 jobs of the
        # same distribution (full lengths), all submitted at once.
        srv = hip.Server()
        cal = serve_jobs(hip, np.random.default_rng(seed + 1), 16, hip.FP32)
        t0 = time.perf_counter()
        for p, st in cal:
            srv.submit(p, st)
        srv.drain(300)
        cap = sum(p.L * p.M * p.N * st for p, st in cal) / (time.perf_counter() - t0)
        srv.close()
        # Stream offering 0.8 of that capacity: jobs are drawn until their work
        # reaches 0.8 x capacity x seconds (job sizes span three decades, so a
        # stream drawn at a fixed rate offers anywhere from ~0.6 to ~1.0 of
        # capacity), and their arrival times are a Poisson process conditioned
        # on that count: sorted uniform times over the window.
        jobs, offered = [], 0.0
        while offered < 0.8 * cap * seconds:
            (p, st), = serve_jobs(hip, rng, 1, hip.FP32)
            jobs.append((p, st))
            offered += p.L * p.M * p.N * st
        arrivals = sorted(rng.uniform(0.0, seconds, len(jobs)).tolist())
        lam = len(jobs) / seconds
        srv = hip.Server()
        ids = []
        start = time.perf_counter()
        for k, (ta, (p, st)) in enumerate(zip(arrivals, jobs)):
            while time.perf_counter() - start < ta:
                time.sleep(min(0.002, max(0.0, ta - (time.perf_counter() - start))))
            ids.append(srv.submit(p, st, outdir=os.path.join(tmp, f"job{k}")))
        srv.drain(600)
        lat, work, t_end = [], 0.0, 0.0
        parts = {"queue": [], "context_creation": [], "stepping": [], "final_files": []}
        failed = 0
        for jid, (p, st) in zip(ids, jobs):
            s = srv.poll(jid)
            if s.state != hip.JOB_DONE:
                failed += 1
                continue
            lat.append(s.t_end - s.t_submit)
            parts["queue"].append(s.t_start - s.t_submit)
            parts["context_creation"].append(s.create_s)
            parts["stepping"].append(s.t_stepped - s.t_start)
            parts["final_files"].append(s.t_end - s.t_stepped)
            work += p.L * p.M * p.N * st
            t_end = max(t_end, s.t_end)
        t_first = srv.poll(ids[0]).t_submit if ids else 0.0
        t_last = srv.poll(ids[-1]).t_submit if ids else 0.0
        info = srv.stats()
        srv.close()
        fp32 = {"jobs": len(ids), "failed": failed, "p50_s": _pct(lat, 50), "p99_s": _pct(lat, 99),
                "served_mpoint_steps_per_s": work / max(t_end - t_first, 1e-9) / 1e6,
                "offered_mpoint_steps_per_s": work / max(t_last - t_first, 1e-9) / 1e6,
                "capacity_mpoint_steps_per_s": cap / 1e6,
                "served_over_capacity": work / max(t_end - t_first, 1e-9) / cap,
                "latency_parts_p50_s": {k: _pct(v, 50) for k, v in parts.items()},
                "latency_parts_p99_s": {k: _pct(v, 99) for k, v in parts.items()},
                "contexts_at_end": info.contexts}
        # Plain FP64 path: the first jobs of the stream, one at a time (GUI queue).
        k64 = min(12, len(jobs))
        lat64, work64, busy64 = [], 0.0, 0.0
        start = time.perf_counter()
        for ta, (p, st) in zip(arrivals[:k64], jobs[:k64]):
            q = hip.default_params(model=p.model, L=p.L, M=p.M, N=p.N, precision=hip.FP64, u0=p.u0, con0=p.con0,
                                   seed=p.seed, grainx=p.grainx, grainy=p.grainy)
            now = time.perf_counter() - start
            if now < ta:
                time.sleep(ta - now)
            t_begin = time.perf_counter() - start
            phi, con = hip.init_fields(q)
            s = hip.Sim(q)
            s.set_fields(phi, con)
            s.step(st)
            s.get_fields()
            s.close()
            done = time.perf_counter() - start
            lat64.append(done - ta)
            work64 += q.L * q.M * q.N * st
            busy64 += done - t_begin
        fp64 = {"jobs": k64, "p50_s": _pct(lat64, 50), "p99_s": _pct(lat64, 99),
                "capacity_mpoint_steps_per_s": work64 / max(busy64, 1e-9) / 1e6,
                "note": "first jobs of the same stream, same arrival times, one job at a time, "
                        "FP64 one-instance contexts without graphs; capacity = work / busy time"}
        return {"workload": f"{seconds:.0f} s stream offering 0.8 x measured FP32 capacity "
                            f"({cap / 1e6:.0f} Mpoint-steps/s): jobs drawn until their work is 0.8 x capacity x "
                            "the window, Poisson arrivals conditioned on that count (sorted uniform times); "
                            "Misfit/CVD 1:1, grids 128^2/256^2/512^2/1024^2/128x128x64, steps U[500,20000], "
                            "u0 U[0.01,0.1], con0 U[0.1,0.4]; latency = arrival to final files written; served = "
                            "work / (last final files - first arrival)", "arrival_rate_per_s": lam, "seed": seed,
                "fp32_batched": fp32, "fp64_plain": fp64}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def _checkpoint_times(path):
    """(step, ms since epoch) pairs of a checkpoint_timestamps.txt."""
    out = []
    for line in open(path).read().splitlines()[1:]:
        t = line.split()
        out.append((int(t[1]), int(t[3])))
    return out


def cvd_checkpoint_leg(hip, steps, mod, L=1024):
    """CVD L x L in FP32 (FP64 k=0 and reductions), u0 0.05, con0 0.25:
    `steps` steps with the energy pair traced every step and the checkpoint
    outputs (VTI, phiball/Phimax, GR, Q6) every `mod`, through the product
    CLI; then the elastic post-processor (Phimax, strain, q4q6) on the GPU over
    every checkpoint's VTI.  The same stepping without checkpoints, through
    the library, separates the checkpoint cost."""
    cli = os.path.join(ROOT, "pfc-exp-gui_amd", "pfc-exp-cli")
    p = hip.default_params(model=hip.CVD, L=L, M=L, N=1, precision=hip.FP32, u0=0.05, con0=0.25)
    phi, con = hip.init_fields(p)
    s = hip.Sim(p)
    s.set_fields(phi, con)
    s.step_energy(mod)
    s.sync()
    t0 = time.perf_counter()
    s.step_energy(mod)
    s.sync()
    per_step = (time.perf_counter() - t0) / mod
    s.close()
    tmp = tempfile.mkdtemp(prefix="pfc_cvd_")
    try:
        run_dir, el_dir = os.path.join(tmp, "run"), os.path.join(tmp, "elastic")
        t0 = time.perf_counter()
        subprocess.run([cli, "--model", "cvd", "--precision", "fp32", "--L", str(L), "--M", str(L), "--u0", "0.05",
                        "--con0", "0.25", "--steps", str(steps), "--mod", str(mod), "--outdir", run_dir],
                       check=True, capture_output=True, timeout=900)
        wall = time.perf_counter() - t0
        ck = _checkpoint_times(os.path.join(run_dir, "checkpoint_timestamps.txt"))
        span_steps = ck[-1][0] - ck[0][0]
        span_s = (ck[-1][1] - ck[0][1]) / 1e3
        n_ck = len(ck)
        t0 = time.perf_counter()
        subprocess.run([cli, "--model", "elastic", "--L", str(L), "--M", str(L), "--input-dir", run_dir,
                        "--outdir", el_dir], check=True, capture_output=True, timeout=900)
        el_wall = time.perf_counter() - t0
        n_el = len([f for f in os.listdir(el_dir) if f.startswith("strain_")])
        atoms = int(open(os.path.join(run_dir, f"Phimax_{ck[-1][0]}.txt")).readline())
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    pts = L * L
    return {
        "workload": f"CVD {L}^2 FP32 + FP64 k=0/reductions, u0 0.05, con0 0.25, {steps} steps, energy every step, "
                    f"checkpoint every {mod} (VTI, phiball/Phimax, GR, Q6), then the elastic post-process (Phimax, "
                    "strain, q4q6) of every checkpoint on the GPU; product CLI",
        "checkpoints": n_ck,
        "atoms_last_checkpoint": atoms,
        "job_wall_s": round(wall, 3),
        "job_mpoint_steps_per_s": round(pts * steps / wall / 1e6, 1),
        "steady_mpoint_steps_per_s": round(pts * span_steps / span_s / 1e6, 1),
        "stepping_only_mpoint_steps_per_s": round(pts / per_step / 1e6, 1),
        "ms_per_step_stepping_only": round(per_step * 1e3, 5),
        "ms_per_checkpoint_outputs": round((span_s - span_steps * per_step) / max(n_ck - 1, 1) * 1e3, 2),
        "elastic_wall_s": round(el_wall, 3),
        "elastic_ms_per_checkpoint": round(el_wall / max(n_el, 1) * 1e3, 2),
        "elastic_checkpoints": n_el,
        "note": "job wall includes process start-up, context creation and tile autotuning; steady = first to last "
                "checkpoint timestamp; checkpoint cost = steady interval minus the same steps without outputs",
    }


def single_instance_leg(hip, steps=200):
    """One instance per context, the GUI's drop-in shapes: Misfit 512^2 FP64
    (the accuracy-study grid, plain steps) and CVD 1024^2 with the energy pair
    traced every step (the checkpoint run's stepping) in FP32 and FP64.  Per
    kernel: the mean of hipExtLaunchKernel start/stop events over `steps`
    launches (the pair also spans 3-5 us of dispatch set-up, so the rate
    understates the kernel's own; the profiler's durations are in profiles/),
    the kernel's algorithmic bytes per launch and that rate against 8 TB/s;
    per step: wall time of `steps` eager steps."""
    cases = [("Misfit 512^2 FP64 (reference narrowing), plain steps", hip.MISFIT, hip.FP64, 512, False),
             ("CVD 1024^2 FP32 + FP64 k=0/reductions, energy every step", hip.CVD, hip.FP32, 1024, True),
             ("CVD 1024^2 FP64 (reference narrowing), energy every step", hip.CVD, hip.FP64, 1024, True)]
    res = []
    for label, model, prec, L, en in cases:
        kw = dict(grainx=L // 32, grainy=L // 32) if model == hip.MISFIT else dict(u0=0.05, con0=0.25)
        p = hip.default_params(model=model, L=L, M=L, precision=prec, **kw)
        phi, con = hip.init_fields(p)
        sim = hip.Sim(p)
        sim.set_fields(phi, con)
        stepf = sim.step_energy if en else sim.step
        stepf(20)
        sim.sync()
        row = {"case": label}
        for kid, name in ((hip.KID_ROW, "row"), (hip.KID_COL, "column")):
            sim.time_kernel(kid, True)
            stepf(steps)
            sim.sync()
            ms, n = sim.kernel_time()
            sim.time_kernel(kid, False)
            by = sim.kernel_bytes(kid)
            row[name] = {"us_per_launch": round(ms * 1e3, 2), "launches": n, "bytes_per_launch": by,
                         "achieved_GBps": round(by / (ms * 1e-3) / 1e9, 1),
                         "frac_of_8TBps": round(by / (ms * 1e-3) / 8e12, 4)}
        t0 = time.perf_counter()
        stepf(steps)
        sim.sync()
        dt = (time.perf_counter() - t0) / steps
        row["us_per_step"] = round(dt * 1e6, 2)
        row["mpoint_steps_per_s"] = round(L * L / dt / 1e6, 1)
        sim.close()
        res.append(row)
    return res


COPY_PROBE_BYTES = 256 * 1024 * 1024


def pmc_drive(hip, steps):
    """A few sweep steps and plain copy kernels of known size (256 MiB read
    and written; linear, and in the column pass's 64-byte pieces in plain and
    in its XCD-aware tile order; each launched twice) for hardware-counter
    runs: the copies calibrate the byte counters."""
    sims = build_sweep(hip, 0)
    for s in sims.values():
        s.step(steps)
        s.sync()
    names = ("linear", "columns_64B", "columns_64B_xcd")
    ms = [hip.copy_probe(COPY_PROBE_BYTES, pat) for pat in (0, 1, 2)]
    print(json.dumps({"pmc_drive": True, "steps": steps, "copy_bytes_each_way": COPY_PROBE_BYTES,
                      "copy_ms": dict(zip(names, ms)),
                      "copy_GBps": {k: 2 * COPY_PROBE_BYTES / m / 1e6 for k, m in zip(names, ms)}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--graph-chunk", type=int, default=25,
                    help="steps per graph launch (the largest divisor of --steps not above this is used)")
    ap.add_argument("--fp64-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=10000)
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--cvd-steps", type=int, default=20000, help="CVD 1024^2 checkpoint run (0: skip)")
    ap.add_argument("--cvd-mod", type=int, default=500)
    ap.add_argument("--pmc-drive", action="store_true")
    ap.add_argument("--single-drive", action="store_true",
                    help="only the single-instance leg (for kernel traces of the drop-in shapes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (there is no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    hip = _load("pfc_hip", os.path.join(ROOT, "pfc-exp-gui_amd", "pfc_hip.py"))
    hip.load()
    if args.pmc_drive:
        pmc_drive(hip, 3)
        return
    if args.single_drive:
        print(json.dumps({"single_instance": single_instance_leg(hip)}))
        return

    # ---- batched FP32 sweep ------------------------------------------------
    # The four grid classes are branches of one hipGraph (chunk steps of each
    # per launch, pfc_group_*); steps beyond whole chunks run per context.
    # The chunk divides the timed steps, so the timed region is whole group
    # launches.  (250-step launches measured the same as 25 after an eager
    # warm-up; only a warm-up through the graph itself made them faster.)
    chunk = max(1, min(args.graph_chunk, args.steps))
    while args.steps % chunk:
        chunk -= 1
    sims = build_sweep(hip, chunk)
    pts = sum(s.batch * s.P for s in sims.values())
    group = hip.Group(list(sims.values()), chunk)

    def sweep(nsteps, grouped):
        if grouped:
            group.step(nsteps // chunk)
            rest = nsteps % chunk
            if rest:
                for s in sims.values():
                    s.step(rest)
            return
        done = 0
        while done < nsteps:
            n = min(chunk, nsteps - done)
            for s in sims.values():
                s.step(n)
            done += n

    # Warm-up: group-graph launches for WARM_S seconds first (from idle the
    # device runs the sweep 10-25% slower until it has been loaded for some
    # tens of ms), then the requested W steps.  Each warm-up launch is waited
    # for before the next is queued: with a backlog of a few hundred queued
    # launches the first graph launch after it ran 2-7% slower (20-step
    # regions, first after the warm-up, seven processes: 70.8-77.9 G queued
    # vs 75.6-80.1 G waited for).
    def warm_group(n):
        for _ in range(n):
            group.step(1)
            for s in sims.values():
                s.sync()

    t_w = time.perf_counter()
    warm_group(1)
    first = time.perf_counter() - t_w
    warm_launches = 1 + max(0, int(WARM_S / max(first, 1e-4)))
    warm_group(warm_launches - 1)
    warm_group(args.warmup // chunk)
    sweep(args.warmup % chunk, True)
    for s in sims.values():
        s.sync()
    t0 = time.perf_counter()
    sweep(args.steps, True)
    for s in sims.values():
        s.sync()
    el = time.perf_counter() - t0
    value = pts * args.steps / el / 1e6
    # Roofline timing: graph-captured launches carry no per-launch events, so
    # the same sweep runs again with per-class graphs on four streams and the
    # 512^2 class's column launches event-timed (that class eager).
    big = sims[512]
    big.time_kernel(hip.KID_COL, True)
    t1 = time.perf_counter()
    sweep(args.steps, False)
    for s in sims.values():
        s.sync()
    el_streams = time.perf_counter() - t1
    col_ms, col_n = big.kernel_time()
    big.time_kernel(hip.KID_COL, False)
    # The same kernel with the GPU to itself (the 512^2 class alone), for
    # comparison with its duration beside the other three classes.
    big.time_kernel(hip.KID_COL, True)
    big.step(100)
    iso_ms, iso_n = big.kernel_time()
    big.time_kernel(hip.KID_COL, False)
    group.close()
    col_bytes = big.kernel_bytes(hip.KID_COL)
    achieved = col_bytes / (col_ms * 1e-3) / 1e9 if col_ms > 0 else 0.0

    # ---- plain FP64 path (one instance at a time, eager) --------------------
    # The same 64 sweep instances run one after another: per-step time of one
    # instance of each grid, weighted by the sweep's 16 instances per grid.
    per_grid = {}
    for g in (128, 256, 384, 512):
        p = hip.default_params(model=hip.MISFIT, L=g, M=g, precision=hip.FP64)
        phi, con = hip.init_fields(p)
        s64 = hip.Sim(p)
        s64.set_fields(phi, con)
        s64.step(10)
        s64.sync()
        t1 = time.perf_counter()
        s64.step(args.fp64_steps)
        s64.sync()
        per_grid[g] = (time.perf_counter() - t1) / args.fp64_steps
        s64.close()
    el64 = 16 * sum(per_grid.values())  # seconds per sweep step
    fp64 = pts / el64 / 1e6
    fp64_512 = 512 * 512 / per_grid[512] / 1e6

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_graph_launches": warm_launches,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: reference initialisers (INITIAL_dis), seeded",
        "config": {
            "workload": "Misfit sweep, 64 instances (4 u0 x 4 con0 x 4 seeds), 16 each of 128^2/256^2/384^2/512^2, "
                        "FP32 + FP64 k=0/reductions; the four grid classes as branches of one hipGraph, "
                        "%d steps per launch" % chunk,
            "points_per_step": pts,
            "col_tiles_autotuned": {str(g): s.col_tiles() for g, s in sims.items()},
        },
        "sweep_per_class_streams": {"value": round(pts * args.steps / el_streams / 1e6, 1), "unit": "Mpoint-steps/s",
                                    "note": "the same sweep with one graph per class on four streams and the 512^2 "
                                            "class eager with event-timed column launches (the roofline run)"},
        "fp64_plain": {"value": round(fp64, 1), "unit": "Mpoint-steps/s",
                       "workload": "the sweep's 64 instances one at a time, FP64 (reference narrowing), eager",
                       "ms_per_sweep_step": round(el64 * 1e3, 4),
                       "ms_per_step_by_grid": {str(g): round(t * 1e3, 4) for g, t in per_grid.items()}},
        "fp64_plain_512": {"value": round(fp64_512, 1), "unit": "Mpoint-steps/s"},
        "single_instance": single_instance_leg(hip),
        "fp32_batched_over_fp64_plain": round(value / fp64, 2),
        "roofline": {
            "bound": "hbm",
            "kernel": "column pass (k_col<float,512>), 16 instances of 512^2",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_launch": col_bytes,
            "launches_timed": col_n,
            "mean_launch_ms": round(col_ms, 5),
            "timing": "hipExtLaunchKernel start/stop events on every launch of the kernel in a second timed run "
                      "of the same sweep (one graph per class on four streams, the 512^2 class eager): "
                      "graph-captured launches carry no per-launch events",
            "mean_launch_ms_alone": round(iso_ms, 5),
            "launches_alone": iso_n,
        },
    }
    traffic_file = os.path.join(ROOT, "profiles", "traffic_k_col_float_512_sweep16.json")
    if os.path.exists(traffic_file):
        tr = json.load(open(traffic_file))
        out["roofline"]["traffic"] = tr.get("bytes_per_dispatch")
        out["roofline"]["traffic_source"] = "profiles/traffic_k_col_float_512_sweep16.json (rocprofv3 --pmc)"
    # The kernel's duration inside the timed region (the group graph) from
    # the committed rocprofv3 kernel trace of this command.
    trace_file = os.path.join(ROOT, "profiles", "k_col_float_512_sweep16_group.json")
    if os.path.exists(trace_file):
        tg = json.load(open(trace_file))
        ms_g = tg.get("mean_ms")
        if ms_g:
            ach_g = col_bytes / (ms_g * 1e-3) / 1e9
            out["roofline"]["timed_region_rocprof"] = {
                "mean_launch_ms": ms_g, "achieved": round(ach_g, 1), "frac": round(ach_g / HBM_PEAK_GBS, 4),
                "source": "profiles/k_col_float_512_sweep16_group.json (rocprofv3 --kernel-trace --stats of "
                          "bench.py: the kernel's launches inside the group graph, where it shares the GPU with "
                          "the other classes' kernels)"}
    # The whole timed region: every kernel of the sweep streams its arrays once
    # (56 B per point-step in FP32: row pass 24, column pass 32), over the
    # region's wall time.
    step_bytes = sum(s.kernel_bytes(hip.KID_ROW) + s.kernel_bytes(hip.KID_COL) for s in sims.values())
    agg = step_bytes * args.steps / el / 1e9
    out["roofline"]["timed_region_all_kernels"] = {
        "bytes_per_step": step_bytes, "achieved": round(agg, 1), "frac": round(agg / HBM_PEAK_GBS, 4),
        "note": "algorithmic bytes of every row and column launch of the timed region (all four classes) over "
                "its wall time: the sweep's aggregate rate against HBM peak"}
    for s in sims.values():
        s.close()
    # The legs below are reported beside the headline, not required for it:
    # a leg that fails records its error and the line is still printed.
    leg_errors = (subprocess.SubprocessError, OSError, ValueError, IndexError, KeyError, RuntimeError)
    if args.accuracy_steps > 0:
        every = max(1, args.accuracy_steps // 10)
        try:
            out["accuracy"] = accuracy_study(hip, args.accuracy_steps, every)
        except leg_errors as e:
            out["accuracy"] = {"error": f"{type(e).__name__}: {e}"}
        print(f"accuracy done: bar_ok={out['accuracy'].get('bar_ok')}", file=sys.stderr, flush=True)
    if args.serve_seconds > 0:
        try:
            out["serving"] = serve_leg(hip, args.serve_seconds)
        except leg_errors as e:
            out["serving"] = {"error": f"{type(e).__name__}: {e}"}
        print(f"serving done: {out['serving'].get('fp32_batched', out['serving'].get('error'))}", file=sys.stderr,
              flush=True)
    if args.cvd_steps > 0:
        try:
            out["cvd_checkpoints"] = cvd_checkpoint_leg(hip, args.cvd_steps, args.cvd_mod)
        except leg_errors as e:
            out["cvd_checkpoints"] = {"error": f"{type(e).__name__}: {e}"}
        print(f"cvd checkpoints done: {out['cvd_checkpoints']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline()
            # one job per host core, 16 at most (the host share of one GPU)
            out["cpu_baseline"]["concurrent"] = cpu_baseline_concurrent(min(16, os.cpu_count() or 1))
        except leg_errors as e:
            out["cpu_baseline"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
