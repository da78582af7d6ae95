"""The FP64 drop-in against the oracle over a long horizon (parity check, run
by hand on an MI355X; not collected by pytest).

For each grid: the FP64 drop-in (reference narrowing) from the reference
initial fields on the GPU, and the oracle (the FP64 CPU restatement, one
process per grid, in parallel) from the same fields; every `every` steps the
field L-infinity and L2 (RMS) distances, the mean-phi difference and the
Phimax error (the oracle's Phimax -- misfit.cpp:3261-3556 or
cvd.cpp:1067-1274 -- on both states: atom counts, atoms matched within half
a lattice spacing, largest matched position and phi differences), the same
for the FP32 path (FP64 k = 0 mode and reductions) from the same fields, beside
the FP64 envelope of the same checkpoint: the largest distance from the
drop-in of the drop-in without the narrowing and of the drop-in from fields
one FP32 round-off away (the reference's rounding of its FFTW transforms is
not reproducible here, so any faithful FP64 build differs from it by about
that much once the dynamics amplify round-off).  Misfit kd = t is the state
after t steps (misfit.cpp:506-512), as in the drop-in.

    python tests/oracle_curves.py [--grids 128,256,512] [--steps 10000] [--every 1000] [--out f.json]

128: Misfit and CVD 128^2 (seconds of oracle time per 1000 steps); 256:
Misfit 256^2 (the GUI's default job grid, 16 x 16 grains) and CVD 256^2
(minutes); 512: Misfit 512^2, the drop-in's benchmark grid (about 10
minutes of one host core for 10^4 steps).  Each oracle process prints one
progress line per checkpoint.
"""
import argparse
import importlib.util
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EPS32 = 2.0 ** -24
GRIDS = {"128": [(0, 128, 128), (1, 128, 128)], "256": [(0, 256, 256), (1, 256, 256)], "512": [(0, 512, 512)]}

ORACLE_RUN = r"""
import importlib.util, json, sys, numpy as np
spec = importlib.util.spec_from_file_location('o', sys.argv[1])
o = importlib.util.module_from_spec(spec); spec.loader.exec_module(o)
kind, L, M, out = int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4]), sys.argv[5]
marks = [int(x) for x in sys.argv[6].split(',')]
phi0 = np.load(sys.argv[7])
s = o.Sim(o.model(kind, L, M, 1, narrow=1, **json.loads(sys.argv[8])), phi0['phi'], phi0['con'])
snaps, t = {}, 0
for mk in marks:
    s.run(mk - t)
    t = mk
    phi, con = s.fields()
    snaps[f'phi_{t}'] = phi; snaps[f'con_{t}'] = con
    print(f'oracle: model {kind} {L}x{M} step {t}', file=sys.stderr, flush=True)
np.savez(out, **snaps)
"""


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def grid_kw(model, L, M):
    """Initialiser arguments: Misfit grains of 16 points (16 x 16 from 256^2
    up, the reference default), CVD u0 0.05, con0 0.25."""
    if model == 1:
        return dict(u0=0.05, con0=0.25)
    if L >= 256:
        return dict(grainx=16, grainy=16)
    return dict(grainx=max(1, L // 16), grainy=max(1, M // 16))


def phimax(orc, model, L, M, p, phi, con):
    """Atoms (x, y, phi) by the oracle's Phimax of the model (dx, u0, con0
    of the run's parameters p)."""
    dx, u0, con0 = p.dx, p.u0, p.con0
    if model == 0:
        x, y, _, ph, _ = orc.phimax_misfit(L, M, 1, dx, con0, phi, con)
    else:
        x, y, _, ph, _, _ = orc.phimax_cvd(L, M, 1, dx, u0, phi)
    return np.asarray(x, float), np.asarray(y, float), np.asarray(ph, float)


def atom_match(a, b, box):
    """Atoms of a matched to their nearest in b (periodic box)."""
    xa, ya, pa = a
    xb, yb, pb = b
    out = {"atoms_dropin": int(len(xa)), "atoms_oracle": int(len(xb))}
    if len(xa) == 0 or len(xb) == 0:
        return out
    best = np.full(len(xa), np.inf)
    idx = np.zeros(len(xa), dtype=np.int64)
    for i0 in range(0, len(xb), 2048):
        ddx = xa[:, None] - xb[None, i0:i0 + 2048]
        ddy = ya[:, None] - yb[None, i0:i0 + 2048]
        ddx -= box * np.round(ddx / box)
        ddy -= box * np.round(ddy / box)
        d = ddx * ddx + ddy * ddy
        j = d.argmin(axis=1)
        dj = d[np.arange(len(xa)), j]
        upd = dj < best
        best[upd] = dj[upd]
        idx[upd] = j[upd] + i0
    matched = np.sqrt(best) < 0.5
    out["atoms_matched"] = int(matched.sum())
    out["max_pos_err"] = float(np.sqrt(best[matched]).max()) if matched.any() else None
    out["max_phi_err"] = float(np.abs(pa[matched] - pb[idx[matched]]).max()) if matched.any() else None
    return out


def gpu_runs(hip, model, L, M, kw, phi, con, marks, fp32=False):
    """The drop-in ('ref'), the drop-in without narrowing ('exact') and the
    drop-in from fields one FP32 round-off away ('pert'), and with fp32 the
    FP32 path ('fp32'): fields at marks."""
    runs = {}
    rng = np.random.default_rng(11)
    pert = (phi * (1.0 + EPS32 * rng.uniform(-1.0, 1.0, phi.size)),
            con * (1.0 + EPS32 * rng.uniform(-1.0, 1.0, con.size)))
    cfg = [("ref", 1, (phi, con), hip.FP64), ("exact", 0, (phi, con), hip.FP64), ("pert", 1, pert, hip.FP64)]
    if fp32:
        cfg.append(("fp32", 1, (phi, con), hip.FP32))
    for tag, narrow, f, prec in cfg:
        p = hip.default_params(model=model, L=L, M=M, precision=prec, **kw)
        p.narrow = narrow
        s = hip.Sim(p)
        try:
            s.set_fields(*f)
            s.set_graph(100)
            snaps, t = {}, 0
            for mk in marks:
                s.step(mk - t)
                t = mk
                snaps[t] = s.get_fields()
        finally:
            s.close()
        runs[tag] = snaps
    return runs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grids", default="128", help="comma-separated: 128, 256, 512")
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--every", type=int, default=1000)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    hip = _load("pfc_hip", os.path.join(ROOT, "pfc-exp-gui_amd", "pfc_hip.py"))
    hip.load()
    orc = _load("pfc_oracle_curves", os.path.join(ROOT, "oracle", "oracle.py"))
    marks = list(range(args.every, args.steps + 1, args.every))
    cases = [c for g in args.grids.split(",") for c in GRIDS[g.strip()]]
    tmp = tempfile.mkdtemp(prefix="pfc_oracle_curves_")
    procs, init = {}, {}
    for c in cases:
        model, L, M = c
        kw = grid_kw(*c)
        p = hip.default_params(model=model, L=L, M=M, precision=hip.FP64, **kw)
        phi, con = hip.init_fields(p)
        init[c] = (phi, con)
        f0 = os.path.join(tmp, f"init_{model}_{L}x{M}.npz")
        np.savez(f0, phi=phi, con=con)
        out = os.path.join(tmp, f"oracle_{model}_{L}x{M}.npz")
        cmd = [sys.executable, "-c", ORACLE_RUN, os.path.join(ROOT, "oracle", "oracle.py"), str(model), str(L),
               str(M), out, ",".join(map(str, marks)), f0, json.dumps(kw)]
        procs[c] = (out, subprocess.Popen(cmd, stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="1")))
    res = {"steps": args.steps, "every": args.every, "cases": {}}
    for c in cases:
        model, L, M = c
        kw = grid_kw(*c)
        name = f"{'misfit' if model == 0 else 'cvd'}_{L}x{M}"
        runs = gpu_runs(hip, model, L, M, kw, *init[c], marks, fp32=True)
        print(f"gpu runs done: {name}", file=sys.stderr, flush=True)
        out, proc = procs[c]
        rc = proc.wait()
        if rc != 0:
            res["cases"][name] = {"error": f"oracle process exit {rc}"}
            continue
        snaps = np.load(out)
        rows = []
        for t in marks:
            rp, rc_ = runs["ref"][t]
            op, oc = snaps[f"phi_{t}"], snaps[f"con_{t}"]
            env = max(max(np.abs(runs[k][t][0] - rp).max(), np.abs(runs[k][t][1] - rc_).max()) for k in ("exact", "pert"))
            row = {"step": t,
                   "linf": float(max(np.abs(rp - op).max(), np.abs(rc_ - oc).max())),
                   "l2_phi": float(np.sqrt(np.mean((rp - op) ** 2))),
                   "l2_con": float(np.sqrt(np.mean((rc_ - oc) ** 2))),
                   "mean_phi_diff": float(rp.mean() - op.mean()),
                   "envelope_linf": float(env)}
            p = hip.default_params(model=model, L=L, M=M, precision=hip.FP64, **kw)
            oatoms = phimax(orc, model, L, M, p, op, oc)
            row["phimax"] = atom_match(phimax(orc, model, L, M, p, rp, rc_), oatoms, L * p.dx)
            fp, fc = runs["fp32"][t]
            row["fp32"] = {"linf": float(max(np.abs(fp - op).max(), np.abs(fc - oc).max())),
                           "l2_phi": float(np.sqrt(np.mean((fp - op) ** 2))),
                           "mean_phi_diff": float(fp.mean() - op.mean()),
                           "phimax": atom_match(phimax(orc, model, L, M, p, fp, fc), oatoms, L * p.dx)}
            rows.append(row)
            print(f"{name} step {t}: linf {row['linf']:.2e} (envelope {env:.2e}) l2 {row['l2_phi']:.2e} "
                  f"atoms {row['phimax']['atoms_dropin']}/{row['phimax']['atoms_oracle']}; FP32 linf "
                  f"{row['fp32']['linf']:.2e} l2 {row['fp32']['l2_phi']:.2e} atoms "
                  f"{row['fp32']['phimax']['atoms_dropin']}", file=sys.stderr, flush=True)
        res["cases"][name] = rows
    shutil.rmtree(tmp, ignore_errors=True)
    text = json.dumps(res, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    else:
        print(text)


if __name__ == "__main__":
    main()
