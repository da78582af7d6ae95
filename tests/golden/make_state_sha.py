"""Write tests/golden/state_sha.json: SHA-256 digests of the state the plain
and energy-traced steps reach on a set of grids, batch sizes and both
precisions, with the energy traces as exact doubles.

The digests pin the bits of the column and row passes as built when the file
was written, so a change to those kernels that is meant to keep their
arithmetic (register use, the order of loads, where coefficients come from)
is checked bit for bit by tests/test_gpu_state_sha.py.  Run on the GPU:

    python tests/golden/make_state_sha.py [out.json]

Each case: the reference initialisers (INITIAL_dis / INITIAL_C2) for the
instance's seed (seed + instance), 20 plain steps, then 5 energy-traced
steps; digests of every instance's (phi, con) as float64 after each phase.
"""
import hashlib
import importlib.util
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

# (model, L, M, N, precision, batch): every column-pass length class of the
# FP32 plain step (two-field 128/256/512/1024, general plan 384/768), 2-D and
# 3-D (both packed slots), and the FP64 field-split and four-field passes.
CASES = [
    ("misfit", 128, 128, 1, "fp32", 2),
    ("misfit", 256, 256, 1, "fp32", 1),
    ("misfit", 384, 384, 1, "fp32", 1),
    ("misfit", 512, 512, 1, "fp32", 2),
    ("misfit", 768, 256, 1, "fp32", 1),
    ("cvd", 1024, 1024, 1, "fp32", 1),
    ("cvd", 64, 1024, 1, "fp32", 1),
    ("misfit", 32, 32, 16, "fp32", 1),
    ("cvd", 64, 64, 32, "fp32", 2),
    ("misfit", 512, 512, 1, "fp64", 1),
    ("cvd", 256, 256, 1, "fp64", 2),
    ("misfit", 384, 384, 1, "fp64", 1),
    ("cvd", 64, 64, 32, "fp64", 1),
]
PLAIN, TRACED = 20, 5


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def case_key(c):
    return "%s_%dx%dx%d_%s_b%d" % c


def params(hip, model, L, M, N, prec, seed_off=0):
    kind = hip.MISFIT if model == "misfit" else hip.CVD
    kw = {}
    if kind == hip.MISFIT:
        kw = dict(grainx=max(1, L // 16), grainy=max(1, M // 16))
    p = hip.default_params(model=kind, L=L, M=M, N=N, precision=hip.FP64 if prec == "fp64" else hip.FP32, **kw)
    p.seed = p.seed + seed_off
    return p


def digest(phi, con):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(phi, np.float64).tobytes())
    h.update(np.ascontiguousarray(con, np.float64).tobytes())
    return h.hexdigest()


def run_case(hip, c):
    model, L, M, N, prec, B = c
    p = params(hip, model, L, M, N, prec)
    sim = hip.Sim(p, batch=B)
    try:
        for i in range(B):
            pi = params(hip, model, L, M, N, prec, seed_off=i)
            phi, con = hip.init_fields(pi)
            if model == "cvd":  # INITIAL_C2 takes no seed: perturb instance i deterministically
                phi = phi + 1e-3 * i * np.cos(np.arange(phi.size) * 0.37)
            sim.set_fields(phi, con, inst=i)
        sim.step(PLAIN)
        plain = [digest(*sim.get_fields(i)) for i in range(B)]
        tr = sim.step_energy(TRACED)
        traced = [digest(*sim.get_fields(i)) for i in range(B)]
    finally:
        sim.close()
    return {"plain": plain, "traced": traced, "energy_trace": [float(x).hex() for x in tr.ravel()]}


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(ROOT, "tests", "golden", "state_sha.json")
    hip = _load("pfc_hip", os.path.join(ROOT, "pfc-exp-gui_amd", "pfc_hip.py"))
    hip.load()
    res = {"plain_steps": PLAIN, "traced_steps": TRACED, "cases": {}}
    for c in CASES:
        print("case", case_key(c), flush=True)
        res["cases"][case_key(c)] = run_case(hip, c)
    with open(out, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print("wrote", out)


if __name__ == "__main__":
    main()
