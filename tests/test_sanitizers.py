"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the CPU code (host
only; nothing here touches a GPU):

* the oracle (oracle/pfc_oracle.c through its reference-contract driver
  oracle/pfc_oracle_cli.c) on the three models: Misfit with checkpoints and
  with the noise term, 2-D and 3-D; CVD with Phimax, GR and Q6_GLO; the
  elastic post-processor (Phimax, elastic3D, cryst_order) on those VTI files.
  elastic.cpp:2869-2882 reads outside its buffer and elastic.cpp:3711-3716
  adds uninitialised memory; the oracle defines both, and these runs show
  that its own code reads and writes only what it allocated and has no
  undefined arithmetic;
* the checkpoint-file writers of pfc-exp-gui_amd/host/pfc_io.cpp
  (tests/host_io_main.cpp), threads included;
and a ThreadSanitizer run of the same writers: the background writer
thread that takes checkpoint tasks in order and the threads that format
per-atom tables in chunks (the reference has no threads; these are this
build's own).
Any sanitizer report aborts the run (-fno-sanitize-recover=all; TSan exits
with a non-zero status on a report).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")


def _have_sanitizers(tmp):
    src = os.path.join(tmp, "probe.c")
    open(src, "w").write("int main(void){return 0;}\n")
    r = subprocess.run(["gcc"] + SAN + ["-o", os.path.join(tmp, "probe"), src], capture_output=True)
    return r.returncode == 0


@pytest.fixture(scope="module")
def san_dir(tmp_path_factory):
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("no host compiler")
    d = str(tmp_path_factory.mktemp("san"))
    if not _have_sanitizers(d):
        pytest.skip("gcc without ASan/UBSan runtime")
    return d


@pytest.fixture(scope="module")
def oracle_cli(san_dir):
    exe = os.path.join(san_dir, "pfc-oracle-cli-san")
    o = os.path.join(ROOT, "oracle")
    subprocess.run(["gcc", "-std=c11", "-ffp-contract=off"] + SAN + ["-o", exe, os.path.join(o, "pfc_oracle_cli.c"),
                    os.path.join(o, "pfc_oracle.c"), "-lm"], check=True)
    return exe


def _run(exe, args, cwd):
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, (args, r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    return r


def test_oracle_misfit_and_elastic_sanitized(oracle_cli, tmp_path):
    src = tmp_path / "misfit"
    _run(oracle_cli, ["--model", "misfit", "--L", "64", "--M", "64", "--grainx", "4", "--grainy", "4",
                      "--steps", "20", "--mod", "10", "--outdir", str(src)], tmp_path)
    assert (src / "Phimax_20.txt").exists()
    out = tmp_path / "elastic"
    _run(oracle_cli, ["--model", "elastic", "--L", "64", "--M", "64", "--input-dir", str(src),
                      "--outdir", str(out)], tmp_path)
    for kd in (10, 20):
        for f in (f"Phimax_{kd}.txt", f"strain_{kd}.txt", f"q4q6_{kd}.txt"):
            assert (out / f).exists(), f


def test_oracle_noise_and_3d_sanitized(oracle_cli, tmp_path):
    _run(oracle_cli, ["--model", "misfit", "--L", "32", "--M", "32", "--grainx", "2", "--grainy", "2",
                      "--axx", "0.01", "--steps", "4", "--mod", "2", "--outdir", str(tmp_path / "noise")], tmp_path)
    _run(oracle_cli, ["--model", "misfit", "--L", "32", "--M", "32", "--N", "16", "--grainx", "2", "--grainy", "2",
                      "--steps", "3", "--mod", "3", "--outdir", str(tmp_path / "m3d")], tmp_path)
    d3 = tmp_path / "c3d"
    _run(oracle_cli, ["--model", "cvd", "--L", "32", "--M", "32", "--N", "16", "--steps", "2", "--mod", "1",
                      "--outdir", str(d3)], tmp_path)
    _run(oracle_cli, ["--model", "elastic", "--L", "32", "--M", "32", "--N", "16", "--input-dir", str(d3),
                      "--outdir", str(tmp_path / "e3d")], tmp_path)


def test_oracle_cvd_sanitized(oracle_cli, tmp_path):
    d = tmp_path / "cvd"
    _run(oracle_cli, ["--model", "cvd", "--L", "64", "--M", "64", "--steps", "21", "--mod", "20",
                      "--outdir", str(d)], tmp_path)
    for f in ("Phimax_0.txt", "phiball_20.cfg", "hist_20.txt", "Grq6_20.txt", "Ql_0.txt", "HistQ_20.txt"):
        assert (d / f).exists(), f


def test_host_io_sanitized(san_dir, tmp_path):
    exe = os.path.join(san_dir, "host_io_main_san")
    subprocess.run(["g++", "-std=c++17", "-pthread"] + SAN + ["-o", exe, os.path.join(ROOT, "tests", "host_io_main.cpp"),
                    os.path.join(ROOT, "pfc-exp-gui_amd", "host", "pfc_io.cpp")], check=True)
    import importlib.util
    spec = importlib.util.spec_from_file_location("t_host_io", os.path.join(ROOT, "tests", "test_host_io.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    d = tmp_path / "io"
    d.mkdir()
    vals, r = mod.run_case(exe, str(d), 5000, 32, 16, 4, 7)
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert np.array_equal(np.fromfile(str(d / "back.bin"), np.float32), vals["field"])  # VTI round trip
    assert open(d / "order.txt").read().split() == [str(k) for k in range(200)]  # writer task order


def test_host_io_thread_sanitized(tmp_path):
    """ThreadSanitizer on the checkpoint-file writers (pfcio::Writer's task
    thread and the chunked table formatting threads): no data race, and the
    files are the sequential formatting's, as in the ASan run."""
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    probe = tmp_path / "tsan_probe.cpp"
    probe.write_text("#include <thread>\nint main(){int x=0; std::thread t([&]{x=1;}); t.join(); return x-1;}\n")
    if subprocess.run(["g++", "-std=c++17", "-pthread", "-fsanitize=thread", "-o", str(tmp_path / "tsan_probe"),
                       str(probe)], capture_output=True).returncode != 0:
        pytest.skip("g++ without the TSan runtime")
    exe = str(tmp_path / "host_io_main_tsan")
    subprocess.run(["g++", "-std=c++17", "-pthread", "-fsanitize=thread", "-g", "-O1", "-o", exe,
                    os.path.join(ROOT, "tests", "host_io_main.cpp"),
                    os.path.join(ROOT, "pfc-exp-gui_amd", "host", "pfc_io.cpp")], check=True)
    import importlib.util
    spec = importlib.util.spec_from_file_location("t_host_io_tsan", os.path.join(ROOT, "tests", "test_host_io.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    d = tmp_path / "io"
    d.mkdir()
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    old = os.environ.get("TSAN_OPTIONS")
    os.environ["TSAN_OPTIONS"] = env["TSAN_OPTIONS"]
    try:
        vals, r = mod.run_case(exe, str(d), 5000, 32, 16, 4, 7)
    finally:
        if old is None:
            os.environ.pop("TSAN_OPTIONS", None)
        else:
            os.environ["TSAN_OPTIONS"] = old
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert open(d / "order.txt").read().split() == [str(k) for k in range(200)]
