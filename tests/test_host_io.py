"""Checkpoint-file writers of pfc-exp-cli (host/pfc_io.cpp), on the CPU.

The per-atom tables are formatted in parallel chunks and the VTI payload is
transposed in blocks; these tests compare every table row with Python's own
sequential printf-style formatting of the same values (both round the exact
binary value), check the VTI write/read round trip and payload order, and
that the background writer runs its tasks in the order they were posted."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "host_io_main.cpp"), os.path.join(ROOT, "pfc-exp-gui_amd", "host", "pfc_io.cpp")]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("hostio") / "host_io_main")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-o", exe] + SRC, check=True)
    return exe


def fixed(v, prec):
    """put_fixed of pfc_io.cpp: NaN prints as -nan (the reference's x86 iostream)."""
    return "-nan" if np.isnan(v) else "%.*f" % (prec, v)


def run_case(driver, d, n, L, M, N, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-5.0, 100.0, n)          # x < 0 rows are skipped by the elastic writers
    y = rng.uniform(-1e6, 1e6, n)
    z = rng.uniform(0.0, 3.0, n)
    v = rng.uniform(-1e-3, 1e-3, n)
    w = rng.uniform(0.0, 1.0, n)
    nb = rng.integers(0, 30, n).astype(np.uint32)
    px, py, pz, pp = (rng.uniform(-0.1, 1.1, n).astype(np.float32) for _ in range(4))
    ql = rng.uniform(0.0, 0.7, n).astype(np.float32)
    ql[::97] = np.nan
    field = rng.standard_normal(L * M * N).astype(np.float32)
    with open(os.path.join(d, "in.bin"), "wb") as f:
        np.array([n, L, M, N], np.int32).tofile(f)
        for a in (x, y, z, v, w):
            a.astype(np.float64).tofile(f)
        nb.tofile(f)
        for a in (px, py, pz, pp, ql):
            a.tofile(f)
        field.tofile(f)
    r = subprocess.run([driver, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return dict(x=x, y=y, z=z, v=v, w=w, nb=nb, px=px, py=py, pz=pz, pp=pp, ql=ql, field=field), r


def lines(path):
    return open(path).read().split("\n")


@pytest.mark.parametrize("n,L,M,N", [(20000, 64, 48, 1), (9000, 16, 8, 5), (1, 4, 4, 1)])
def test_tables_match_sequential_format(driver, tmp_path, n, L, M, N):
    d = str(tmp_path)
    a, r = run_case(driver, d, n, L, M, N, seed=n)
    keep = a["x"] >= 0
    idx = np.nonzero(keep)[0]
    exp = [f"{n}", "Kommentar"] + ["%.4f %.4f %.4f %.4f" % (a["x"][i], a["y"][i], a["z"][i], a["v"][i]) for i in idx] + [""]
    assert lines(os.path.join(d, "pm.txt")) == exp
    exp = [f"{n}", "Kommentar"] + ["st %.4f %.4f %.4f %.4f %d" % (a["x"][i], a["y"][i], a["z"][i], a["v"][i], a["nb"][i])
                                    for i in idx] + [""]
    assert lines(os.path.join(d, "st.txt")) == exp
    exp = [f"{len(idx)}", "Kommentar"] + ["qc %.4f %.4f %.4f %.4f %.4f" % (a["x"][i], a["y"][i], a["z"][i], a["v"][i],
                                                                        a["w"][i]) for i in idx] + [""]
    assert lines(os.path.join(d, "q.txt")) == exp
    assert f"q4q6_7.txt: atoms={len(idx)}" in r.stderr
    # CVD atoms: phiball rows after its 15 header lines, Phimax rows (float products), phimax.txt line
    cfg = lines(os.path.join(d, "phiball_7.cfg"))
    assert cfg[0] == f"Number of particles = {n}" and len(cfg) == 15 + n + 1
    assert cfg[15:] == ["%.4f %.4f %.4f" % (a["px"][i], a["py"][i], a["pz"][i]) for i in range(n)] + [""]
    fx, fy, fz = np.float32(L * 0.5), np.float32(M * 0.5), np.float32(N * 0.5)
    exp = [f"{n}", "Kommentar"] + ["%.4f %.4f %.4f %.4f" % (a["pz"][i] * fx, a["px"][i] * fy, a["py"][i] * fz, a["pp"][i])
                                    for i in range(n)] + [""]
    assert lines(os.path.join(d, "Phimax_7.txt")) == exp
    assert lines(os.path.join(d, "phimax.txt"))[0] == "7 , 0.3 , 0.8 , 0"
    exp = ["%d %d %s %s" % (i, a["nb"][i], fixed(float(a["ql"][i]), 6), fixed(float(np.float32(0.25)), 6))
           for i in range(n)] + [""]
    assert lines(os.path.join(d, "Ql_7.txt")) == exp
    hq = lines(os.path.join(d, "HistQ_7.txt"))
    assert hq[:30] == ["%.4f %d" % (float(np.float32(b + 1) * np.float32(0.05)), 1000 * b + 7) for b in range(30)]


@pytest.mark.parametrize("L,M,N", [(64, 48, 1), (16, 8, 5), (33, 17, 3)])
def test_vti_round_trip_and_order(driver, tmp_path, L, M, N):
    d = str(tmp_path)
    a, r = run_case(driver, d, 1, L, M, N, seed=L * M * N)
    P = L * M * N
    assert r.stdout.split() == ["read", str(P)]
    back = np.fromfile(os.path.join(d, "back.bin"), np.float32)
    assert np.array_equal(back, a["field"])
    raw = open(os.path.join(d, "f.vti"), "rb").read()
    i = raw.index(b"\n_") + 2
    assert int(np.frombuffer(raw[i:i + 4], np.int32)[0]) == 4 * P + 4
    pay = np.frombuffer(raw[i + 4:i + 4 + 4 * P], np.float32)
    # field index n + N*(m + M*l); payload order n, m, l with l fastest
    f3 = a["field"].reshape(L, M, N)
    assert np.array_equal(pay, f3.transpose(2, 1, 0).ravel())
    assert raw[i + 4 + 4 * P:] == b"\n</AppendedData>\n</VTKFile>\n"


@pytest.mark.parametrize("L,M,N", [(64, 48, 1), (16, 8, 5)])
def test_vti_short_file(driver, tmp_path, L, M, N):
    """READPHIVTIN on a truncated file: the values present fill the field in
    payload order (n, m, l with l fastest), the rest are left untouched."""
    d = str(tmp_path)
    P = L * M * N
    k = P // 3 + 5
    vals = np.arange(k, dtype=np.float32)
    with open(os.path.join(d, "short.vti"), "wb") as f:
        f.write(b"x\n" * 12 + b"_" + np.int32(4 * P + 4).tobytes() + vals.tobytes())
    a, r = run_case(driver, d, 1, L, M, N, seed=5)
    assert f"short {k}" in r.stdout.split("\n")
    got = np.fromfile(os.path.join(d, "short_back.bin"), np.float32).reshape(L, M, N)
    exp = np.full(P, -7.0, np.float32)
    pay = exp.copy()
    pay[:k] = vals                                   # payload order: index (n*M + m)*L + l
    exp = pay.reshape(N, M, L).transpose(2, 1, 0)    # -> field index n + N*(m + M*l)
    assert np.array_equal(got, exp)


def test_writer_keeps_post_order(driver, tmp_path):
    d = str(tmp_path)
    run_case(driver, d, 1, 4, 4, 1, seed=3)
    assert lines(os.path.join(d, "order.txt")) == [str(k) for k in range(200)] + [""]
