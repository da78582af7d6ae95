"""End-to-end: the product pfc-exp-cli (GPU) against the oracle CLI (CPU) on
GUI-style argv (MainWindow.cpp:2401-2428 plus the runtime-grid flags).
Same file names, byte-identical run_config / PVTI / initial VTI / VTI
headers, field payloads within FP64 round-off, energy.txt within its printed
precision, Phimax atoms matched, and the elastic post-processor byte-identical
on the same VTI input (integer/index work and unfused IEEE arithmetic)."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pfc-exp-gui_amd", "pfc-exp-cli")
OCLI = os.path.join(ROOT, "oracle", "pfc-oracle-cli")
PROGRESS = re.compile(r"^PFC_PROGRESS\s+step=(\d+)\s+total=(\d+)\s*$")


def run(exe, args, cwd):
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (exe, args, r.stderr[-2000:])
    return r


def vti_split(path):
    raw = open(path, "rb").read()
    i = raw.index(b"\n_") + 2
    n = int(np.frombuffer(raw[i:i + 4], np.int32)[0])
    return raw[:i + 4], np.frombuffer(raw[i + 4:i + n], np.float32), raw[i + n:], n


def read_table(path):
    lines = open(path).read().splitlines()
    return int(lines[0]), lines[1], np.array([[float(t) for t in l.split() if t != "st"] for l in lines[2:]])


def test_misfit_cli_matches_oracle(tmp_path):
    args = ["--u0", "0.05", "--con0", "0.2", "--sig", "0.05", "--dt", "0.05", "--dx", "0.125", "--steps", "50",
            "--mod", "25", "--seed", "20200604", "--L", "64", "--M", "64", "--grainx", "4", "--grainy", "4"]
    g, o = tmp_path / "gpu", tmp_path / "cpu"
    r = run(CLI, ["--model", "misfit"] + args + ["--outdir", str(g)], tmp_path)
    run(OCLI, ["--model", "misfit"] + args + ["--outdir", str(o)], tmp_path)
    steps = [int(PROGRESS.match(l).group(1)) for l in r.stdout.splitlines() if PROGRESS.match(l)]
    assert steps[0] == 1 and steps[-1] == 50
    assert sorted(os.listdir(g)) == sorted(os.listdir(o))
    for f in ("run_config.txt", "phi_0_0.vti", "con_0_0.vti", "phi_25.pvti", "con_50.pvti"):
        assert (g / f).read_bytes() == (o / f).read_bytes(), f
    for kd in (25, 50):
        for fld in ("phi", "con"):
            hg, pg, tg, ng = vti_split(g / f"{fld}_{kd}_0.vti")
            ho, po, to, no = vti_split(o / f"{fld}_{kd}_0.vti")
            assert hg == ho and tg == to and ng == no
            assert np.abs(pg - po).max() < 1e-6
        cg, _, ag = read_table(g / f"Phimax_{kd}.txt")
        co, _, ao = read_table(o / f"Phimax_{kd}.txt")
        assert cg == co
        assert np.abs(ag - ao).max() <= 2e-4
    eg = np.loadtxt(g / "energy.txt")
    eo = np.loadtxt(o / "energy.txt")
    assert np.array_equal(eg[:, 0], eo[:, 0])
    assert np.allclose(eg[:, 1:], eo[:, 1:], rtol=2e-5, atol=1e-9)
    ck = [l.split()[1] for l in (g / "checkpoint_timestamps.txt").read_text().splitlines()[1:]]
    assert ck == ["25", "50"]


def test_gui_default_job_matches_oracle(tmp_path):
    """The GUI's default job (MainWindow.cpp:2401-2428 argv with its default
    values): Misfit 256^2, 200 steps, a checkpoint every 25, seed 20200604,
    16 x 16 grains, the FP64 drop-in, against the oracle CLI: the same file
    set, run_config / PVTI / initial VTI and every VTI header byte-identical,
    payloads (float32) within FP64-narrowing round-off, Phimax counts equal
    and rows within the printed 4 decimals, energy.txt within its 6 digits."""
    args = ["--u0", "0.05", "--con0", "0.2", "--sig", "0.05", "--dt", "0.05", "--dx", "0.125", "--steps", "200",
            "--mod", "25", "--seed", "20200604"]
    g, o = tmp_path / "gpu", tmp_path / "cpu"
    r = run(CLI, ["--model", "misfit"] + args + ["--outdir", str(g)], tmp_path)
    run(OCLI, ["--model", "misfit"] + args + ["--outdir", str(o)], tmp_path)
    steps = [int(PROGRESS.match(l).group(1)) for l in r.stdout.splitlines() if PROGRESS.match(l)]
    assert steps[0] == 1 and steps[-1] == 200 and len(steps) == 200
    assert sorted(os.listdir(g)) == sorted(os.listdir(o))
    for f in ("run_config.txt", "phi_0_0.vti", "con_0_0.vti", "phi_0.pvti", "phi_100.pvti", "con_200.pvti"):
        assert (g / f).read_bytes() == (o / f).read_bytes(), f
    worst = 0.0
    for kd in range(25, 201, 25):
        for fld in ("phi", "con"):
            hg, pg, tg, ng = vti_split(g / f"{fld}_{kd}_0.vti")
            ho, po, to, no = vti_split(o / f"{fld}_{kd}_0.vti")
            assert hg == ho and tg == to and ng == no
            worst = max(worst, float(np.abs(pg - po).max()))
        cg, _, ag = read_table(g / f"Phimax_{kd}.txt")
        co, _, ao = read_table(o / f"Phimax_{kd}.txt")
        assert cg == co, kd
        assert np.abs(ag - ao).max() <= 2e-4, kd
    print(f"GUI default job: max |VTI payload GPU - oracle| = {worst:.3e}", flush=True)
    assert worst < 1e-5
    eg = np.loadtxt(g / "energy.txt")
    eo = np.loadtxt(o / "energy.txt")
    assert np.array_equal(eg[:, 0], eo[:, 0]) and list(eg[:, 0]) == list(range(25, 201, 25))
    assert np.allclose(eg[:, 1:], eo[:, 1:], rtol=2e-5, atol=1e-9)
    ck = [l.split()[1] for l in (g / "checkpoint_timestamps.txt").read_text().splitlines()[1:]]
    assert ck == [str(k) for k in range(25, 201, 25)]


def test_cvd_cli_matches_oracle(tmp_path):
    args = ["--steps", "21", "--mod", "10", "--L", "64", "--M", "64", "--u0", "0.05", "--con0", "0.25"]
    g, o = tmp_path / "gpu", tmp_path / "cpu"
    r = run(CLI, ["--model", "cvd"] + args + ["--outdir", str(g)], tmp_path)
    run(OCLI, ["--model", "cvd"] + args + ["--outdir", str(o)], tmp_path)
    prog = [l for l in r.stderr.splitlines() if PROGRESS.match(l)]
    assert len(prog) == 21
    assert sorted(os.listdir(g)) == sorted(os.listdir(o))
    eg = np.loadtxt(g / "energy.txt")
    eo = np.loadtxt(o / "energy.txt")
    assert eg.shape == (21, 3) and np.array_equal(eg[:, 0], np.arange(21))
    assert np.allclose(eg[:, 1:], eo[:, 1:], rtol=2e-5, atol=1e-9)
    for kd in (0, 10, 20):
        cg, _, ag = read_table(g / f"Phimax_{kd}.txt")
        co, _, ao = read_table(o / f"Phimax_{kd}.txt")
        assert cg == co and np.abs(ag - ao).max() <= 2e-4
        assert (g / f"phiball_{kd}.cfg").read_text().splitlines()[:15] == \
            (o / f"phiball_{kd}.cfg").read_text().splitlines()[:15]
    for kd in (0, 10, 20):   # GR every checkpoint: integer histogram, host-side normalisation
        assert (g / f"hist_{kd}.txt").read_bytes() == (o / f"hist_{kd}.txt").read_bytes()
        assert (g / f"Grq6_{kd}.txt").read_bytes() == (o / f"Grq6_{kd}.txt").read_bytes()
    for kd in (0, 20):       # Q6_GLO when kd % 20 == 0
        qg, qo = np.loadtxt(g / f"Ql_{kd}.txt", ndmin=2), np.loadtxt(o / f"Ql_{kd}.txt", ndmin=2)
        assert qg.shape == qo.shape and np.array_equal(qg[:, :2], qo[:, :2])
        assert np.abs(qg[:, 2:] - qo[:, 2:]).max() <= 3e-6
        hg, ho = np.loadtxt(g / f"HistQ_{kd}.txt"), np.loadtxt(o / f"HistQ_{kd}.txt")
        assert np.array_equal(hg[:, 0], ho[:, 0]) and np.abs(hg[:, 1] - ho[:, 1]).sum() <= 2
    assert not (g / "Ql_10.txt").exists()
    pg = (g / "phimax.txt").read_text().splitlines()
    po = (o / "phimax.txt").read_text().splitlines()
    assert len(pg) == len(po) == 18
    assert [l for l in pg if " , " not in l] == [l for l in po if " , " not in l]


def test_elastic_cli_bit_identical(tmp_path):
    src = tmp_path / "src"
    run(OCLI, ["--model", "misfit", "--L", "128", "--M", "128", "--grainx", "4", "--grainy", "4", "--steps", "40",
               "--mod", "20", "--outdir", str(src)], tmp_path)
    g, o = tmp_path / "g", tmp_path / "o"
    common = ["--model", "elastic", "--L", "128", "--M", "128", "--input-dir", str(src)]
    r = run(CLI, common + ["--outdir", str(g)], tmp_path)
    run(OCLI, common + ["--outdir", str(o)], tmp_path)
    assert "PFC_PROGRESS step=2 total=2" in r.stderr     # steps 20, 40 (0 dropped)
    for kd in (20, 40):
        for f in (f"Phimax_{kd}.txt", f"strain_{kd}.txt"):
            assert (g / f).read_bytes() == (o / f).read_bytes(), f
    assert not (g / "Phimax_0.txt").exists()
    for kd in (20, 40):
        a, b = (g / f"q4q6_{kd}.txt").read_text().splitlines(), (o / f"q4q6_{kd}.txt").read_text().splitlines()
        assert a[:2] == b[:2] and len(a) == len(b) > 2
        ta = np.array([[float(t) for t in l.split()[1:]] for l in a[2:]])
        tb = np.array([[float(t) for t in l.split()[1:]] for l in b[2:]])
        assert np.array_equal(ta[:, :3], tb[:, :3])
        assert np.abs(ta[:, 3:] - tb[:, 3:]).max() <= 1.5e-4   # one unit of the 4th decimal
        assert f"q4q6_{kd}.txt: atoms={a[0]}" in r.stderr


def test_cli_graph_steps_match_eager(tmp_path):
    """With 1600 steps the CLI prints progress every 8 steps and runs those 8
    as one captured graph; its checkpoint fields equal the same steps taken
    eagerly through the library, bit for bit."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("pfc_hip_cli", os.path.join(ROOT, "pfc-exp-gui_amd", "pfc_hip.py"))
    hip = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(hip)
    g = tmp_path / "g"
    r = run(CLI, ["--model", "misfit", "--L", "128", "--M", "128", "--grainx", "4", "--grainy", "4", "--steps", "1600",
                  "--mod", "800", "--outdir", str(g)], tmp_path)
    steps = [int(PROGRESS.match(l).group(1)) for l in r.stdout.splitlines() if PROGRESS.match(l)]
    assert steps[:3] == [1, 8, 16] and steps[-1] == 1600
    p = hip.default_params(model=hip.MISFIT, L=128, M=128, precision=hip.FP64, grainx=4, grainy=4)
    phi, con = hip.init_fields(p)
    s = hip.Sim(p)
    s.set_fields(phi, con)
    s.step(1600)
    ephi, econ = s.get_fields()
    s.close()
    for name, ref in (("phi", ephi), ("con", econ)):
        _, pay, _, _ = vti_split(g / f"{name}_1600_0.vti")
        # payload order n, m, l (l fastest) of the field index n + N*(m + M*l)
        assert np.array_equal(pay, ref.astype(np.float32).reshape(128, 128).T.ravel()), name


def test_fp32_cli_runs(tmp_path):
    g = tmp_path / "g"
    run(CLI, ["--model", "misfit", "--precision", "fp32", "--L", "128", "--M", "128", "--steps", "30",
              "--mod", "15", "--grainx", "8", "--grainy", "8", "--outdir", str(g)], tmp_path)
    assert (g / "Phimax_30.txt").exists() and (g / "phi_30_0.vti").exists()
    e = np.loadtxt(g / "energy.txt")
    assert e.shape == (2, 3) and np.all(np.isfinite(e))


@pytest.mark.parametrize("model,extra", [
    ("misfit", ["--grainx", "2", "--grainy", "2"]),
    ("misfit", ["--grainx", "2", "--grainy", "2", "--axx", "0.001"]),
    ("cvd", []),
])
def test_3d_cli_matches_oracle(tmp_path, model, extra):
    """3-D grids (N > 1) and the noise term through the process boundary."""
    args = ["--steps", "12", "--mod", "6", "--L", "32", "--M", "32", "--N", "16", "--u0", "0.05",
            "--con0", "0.25"] + extra
    g, o = tmp_path / "gpu", tmp_path / "cpu"
    run(CLI, ["--model", model] + args + ["--outdir", str(g)], tmp_path)
    run(OCLI, ["--model", model] + args + ["--outdir", str(o)], tmp_path)
    assert sorted(os.listdir(g)) == sorted(os.listdir(o))
    assert (g / "run_config.txt").read_bytes() == (o / "run_config.txt").read_bytes()
    kd = 12 if model == "misfit" else 6  # misfit kd = t+1, cvd kd = t (t % mod == 0)
    for fld in ("phi", "con"):
        hg, pg, tg, ng = vti_split(g / f"{fld}_{kd}_0.vti")
        ho, po, to, no = vti_split(o / f"{fld}_{kd}_0.vti")
        assert hg == ho and tg == to and ng == no == 4 * 32 * 32 * 16 + 4  # int32 word = payload + 4
        assert np.abs(pg - po).max() < 1e-5
    eg = np.loadtxt(g / "energy.txt")
    eo = np.loadtxt(o / "energy.txt")
    assert np.array_equal(eg[:, 0], eo[:, 0])
    assert np.allclose(eg[:, 1:], eo[:, 1:], rtol=2e-5, atol=1e-9)


def test_cvd_1024_checkpoints_elastic(tmp_path):
    """CVD 1024^2 in FP32 (energy every step, checkpoints every 500 steps),
    then the GPU elastic post-processor and the oracle's on the same VTI
    files: Phimax and strain byte-identical, q4q6 within its last decimal."""
    src = tmp_path / "src"
    r = run(CLI, ["--model", "cvd", "--precision", "fp32", "--L", "1024", "--M", "1024", "--u0", "0.05",
                  "--con0", "0.25", "--steps", "1001", "--mod", "500", "--outdir", str(src)], tmp_path)
    assert len([l for l in r.stderr.splitlines() if PROGRESS.match(l)]) == 1001
    e = np.loadtxt(src / "energy.txt")
    assert e.shape == (1001, 3) and np.array_equal(e[:, 0], np.arange(1001)) and np.all(np.isfinite(e))
    ck = [int(l.split()[1]) for l in (src / "checkpoint_timestamps.txt").read_text().splitlines()[1:]]
    assert ck == [0, 500, 1000]
    for kd in ck:
        for f in (f"phi_{kd}_0.vti", f"Phimax_{kd}.txt", f"phiball_{kd}.cfg", f"hist_{kd}.txt", f"Ql_{kd}.txt"):
            assert (src / f).exists(), f
    g, o = tmp_path / "g", tmp_path / "o"
    common = ["--model", "elastic", "--L", "1024", "--M", "1024", "--input-dir", str(src)]
    run(CLI, common + ["--outdir", str(g)], tmp_path)
    run(OCLI, common + ["--outdir", str(o)], tmp_path)
    for kd in (500, 1000):
        for f in (f"Phimax_{kd}.txt", f"strain_{kd}.txt"):
            assert (g / f).read_bytes() == (o / f).read_bytes(), f
        assert int((g / f"Phimax_{kd}.txt").read_text().split("\n", 1)[0]) > 10000
        a, b = (g / f"q4q6_{kd}.txt").read_text().splitlines(), (o / f"q4q6_{kd}.txt").read_text().splitlines()
        assert a[:2] == b[:2] and len(a) == len(b) > 2
        ta = np.array([[float(t) for t in l.split()[1:]] for l in a[2:]])
        tb = np.array([[float(t) for t in l.split()[1:]] for l in b[2:]])
        assert np.array_equal(ta[:, :3], tb[:, :3])
        assert np.abs(ta[:, 3:] - tb[:, 3:]).max() <= 1.5e-4
