"""CPU tests of the product's host layer: the C ABI library loads and exports
every symbol include/pfc_hip.h declares, the host-side reference
initialisers equal the oracle's bit for bit, and the CLI honours the
reference's argument / exit-code contract on paths that end before any GPU
call (main_cli.cpp:24-88, misfit.cpp:181-309, elastic.cpp:188-289,
pfc_common.h:249-269)."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pfc-exp-gui_amd", "pfc-exp-cli")
HEADER = os.path.join(ROOT, "include", "pfc_hip.h")


def test_library_exports_header_symbols(hip):
    text = open(HEADER).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(pfc_\w+)\s*\(", text, re.M))
    assert len(declared) >= 19
    lib = hip.load()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(hip.SYMBOLS)


def test_params_defaults(hip):
    p = hip.default_params(hip.MISFIT)
    assert (p.u0, p.con0, p.sig, p.dt, p.dx) == tuple(float(np.float32(v)) for v in (0.05, 0.2, 0.05, 0.05, 0.125))
    assert (p.grainx, p.grainy, p.grainz, p.seed, p.narrow) == (16, 16, 1, 20200604, 1)
    q = hip.default_params(hip.CVD)
    assert (q.u0, q.sig) == (float(np.float32(0.01)), float(np.float32(0.03)))


@pytest.mark.parametrize("kind,L,M,kw", [
    (0, 256, 256, {}),
    (0, 128, 384, dict(grainx=4, grainy=8, seed=7, u0=0.04, con0=0.3)),
    (0, 64, 64, dict(grainx=4, grainy=4, seed=123456)),
    (1, 128, 64, dict(u0=0.05, con0=0.25)),
])
def test_init_fields_match_oracle(hip, oracle, kind, L, M, kw):
    p = hip.default_params(model=kind, L=L, M=M, **kw)
    m = oracle.model(kind, L, M, 1, **kw)
    phi, con = hip.init_fields(p)
    ophi, ocon = oracle.init_fields(m)
    assert np.array_equal(phi, ophi) and np.array_equal(con, ocon)


def _cli(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_cli_global_help(tmp_path):
    r = _cli(["--help"], tmp_path)
    assert r.returncode == 0 and "Models:" in r.stderr


@pytest.mark.parametrize("args", [
    ["--model", "nosuch"],
    ["--model"],
    ["--model", "misfit", "--bogus", "1"],
    ["--steps", "0"],
    ["--dt=-1"],
    ["--u0", "abc"],
    ["--steps"],
    ["--grainx", "512"],
    ["--mod", "0"],
])
def test_cli_bad_arguments_exit_2(tmp_path, args):
    r = _cli(args, tmp_path)
    assert r.returncode == 2, (args, r.stderr)
    assert not (tmp_path / "run_config.txt").exists()


def test_cli_model_help(tmp_path):
    r = _cli(["--model", "cvd", "--help"], tmp_path)
    assert r.returncode == 0 and "--model cvd" in r.stderr


def test_cli_outdir_failure(tmp_path):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    r = _cli(["--outdir", str(blocker / "sub")], tmp_path)
    assert r.returncode == 2 and "outdir" in r.stderr


def test_cli_elastic_ignores_unknown_and_missing_input(tmp_path):
    """elastic ignores unknown flags (elastic.cpp:279); with no checkpoint
    files it falls back to mod-spaced steps and exits 2 on the missing VTI
    (elastic.cpp:2572-2575) after writing run_config.txt."""
    out = tmp_path / "out"
    r = _cli(["--model", "elastic", "--whatever", "3", "--input-dir", str(tmp_path / "empty"),
              "--outdir", str(out), "--steps", "400", "--mod", "200", "--benchel", "1.5"], tmp_path)
    assert r.returncode == 2 and "can't open input file" in r.stderr
    cfg = (out / "run_config.txt").read_text().splitlines()
    assert cfg[0] == "model elastic" and "benchel 1.5" in cfg
    assert (out / "checkpoint_timestamps.txt").read_text().startswith("# Checkpoint timestamps")
    assert "PFC_PROGRESS step=1 total=2" in r.stderr


def test_server_admission_limits(hip):
    """Request-loop host logic that ends before any device work: a job larger
    than the HBM budget fails at admission, bad arguments are refused, and
    destroying the server with nothing running returns promptly."""
    srv = hip.Server(hbm_budget=1e6)
    p = hip.default_params(model=hip.MISFIT, L=64, M=64, grainx=4, grainy=4)
    jid = srv.submit(p, 10)
    srv.drain(30)
    st = srv.poll(jid)
    assert st.state == hip.JOB_FAILED and "budget" in srv.error(jid)
    assert st.t_end >= st.t_submit
    with pytest.raises(hip.PfcError):
        srv.submit(p, 0)
    with pytest.raises(hip.PfcError):
        srv.poll(jid + 100)
    with pytest.raises(hip.PfcError):
        srv.result(jid)
    info = srv.stats()
    assert info.jobs_done == 0 and info.jobs_unfinished == 0 and info.budget == 1e6
    srv.close()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message before
    loading the library or touching a device."""
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_state_digest_fixture_covers_every_case():
    """tests/golden/state_sha.json holds a digest set for every case the GPU
    bit-pin test runs (tests/test_gpu_state_sha.py), with one plain and one
    energy-traced digest per instance and one (energy, density) pair per
    instance and traced step."""
    import importlib.util
    import json

    spec = importlib.util.spec_from_file_location("make_state_sha_cpu",
                                                  os.path.join(ROOT, "tests", "golden", "make_state_sha.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "state_sha.json")))
    assert gold["plain_steps"] == gen.PLAIN and gold["traced_steps"] == gen.TRACED
    assert set(gold["cases"]) == {gen.case_key(c) for c in gen.CASES}
    for c in gen.CASES:
        g = gold["cases"][gen.case_key(c)]
        B = c[5]
        assert len(g["plain"]) == B and len(g["traced"]) == B
        assert all(re.fullmatch(r"[0-9a-f]{64}", h) for h in g["plain"] + g["traced"])
        assert len(g["energy_trace"]) == 2 * B * gen.TRACED
        vals = [float.fromhex(x) for x in g["energy_trace"]]
        assert all(np.isfinite(vals))


def test_oracle_curves_atom_match():
    """The atom matcher of tests/oracle_curves.py (nearest neighbour in a
    periodic box, matched within half a lattice spacing): a set against
    itself shifted by small known offsets, one atom moved across the box
    edge, one atom far from any other."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("oracle_curves_cpu", os.path.join(ROOT, "tests", "oracle_curves.py"))
    oc = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(oc)
    box = 32.0
    g = np.arange(0.5, box, 1.0)
    x, y = [a.ravel() for a in np.meshgrid(g, g)]
    p = np.linspace(0.1, 0.9, x.size)
    rng = np.random.default_rng(3)
    d = rng.uniform(-0.05, 0.05, (2, x.size))
    xb, yb = x + d[0], y + d[1]
    xb[0] = x[0] - 0.02 + box  # wraps: still 0.02 away from its partner
    out = oc.atom_match((x, y, p), (xb, yb, p + 1e-3), box)
    assert out["atoms_dropin"] == out["atoms_oracle"] == out["atoms_matched"] == x.size
    ddx = d[0].copy()
    ddx[0] = -0.02
    want = np.sqrt(ddx ** 2 + d[1] ** 2).max()
    assert abs(out["max_pos_err"] - want) < 1e-9
    assert abs(out["max_phi_err"] - 1e-3) < 1e-12
    # an extra atom halfway between lattice sites is left unmatched
    xa2, ya2, pa2 = np.append(x, 5.0), np.append(y, 5.0), np.append(p, 0.5)
    out2 = oc.atom_match((xa2, ya2, pa2), (x, y, p), box)
    assert out2["atoms_dropin"] == x.size + 1 and out2["atoms_matched"] == x.size
    assert oc.grid_kw(0, 256, 256) == dict(grainx=16, grainy=16)
    assert oc.grid_kw(0, 128, 128) == dict(grainx=8, grainy=8)
    assert oc.grid_kw(1, 256, 256) == dict(u0=0.05, con0=0.25)
