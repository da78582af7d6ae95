"""Request loop with continuous batching on the GPU.

Jobs that arrive while others are stepping are admitted into free slots of
their class's batch context (idle slots masked); each result is the same as
the job run alone in a one-instance context of the same precision -- bit
for bit, since every instance runs the same kernels on its own data.
"""
import os
import shutil
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _alone(hip, p, steps):
    s = hip.Sim(p)
    phi, con = hip.init_fields(p)
    s.set_fields(phi, con)
    s.set_seed(p.seed)
    s.step(steps)
    out = s.get_fields()
    e = s.energy()[0]
    s.close()
    return out, e


def _jobs(hip):
    jobs = []
    for i in range(5):  # one class: Misfit FP32 64x64, different seeds / u0 / lengths
        p = hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4,
                               seed=500 + i, u0=0.02 + 0.01 * i)
        jobs.append((p, 40 + 17 * i))
    jobs.append((hip.default_params(model=hip.CVD, L=64, M=32, precision=hip.FP32, u0=0.05, con0=0.25), 60))
    jobs.append((hip.default_params(model=hip.MISFIT, L=32, M=32, N=16, precision=hip.FP64, grainx=2, grainy=2,
                                    seed=9), 30))
    jobs.append((hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4,
                                    seed=77, axx=0.001), 45))  # noise term on
    return jobs


def test_mix_equals_alone(hip):
    srv = hip.Server(max_slots=4, graph_chunk=8, quantum=16)
    jobs = _jobs(hip)
    ids = []
    for k, (p, steps) in enumerate(jobs):
        ids.append(srv.submit(p, steps, keep_fields=True))
        if k % 3 == 2:
            time.sleep(0.05)  # later jobs arrive while earlier ones step
    srv.drain(120)
    info = srv.stats()
    print(f"contexts={info.contexts} done={info.jobs_done} point_steps={info.point_steps:.3e}", flush=True)
    assert info.jobs_done == len(jobs)
    for jid, (p, steps) in zip(ids, jobs):
        st = srv.poll(jid)
        assert st.state == hip.JOB_DONE, srv.error(jid)
        assert st.steps_done == steps and st.t_submit <= st.t_start <= st.t_end
        phi, con = srv.result(jid)
        (aphi, acon), e = _alone(hip, p, steps)
        d = max(np.abs(phi - aphi).max(), np.abs(con - acon).max())
        print(f"job {jid} L={p.L} M={p.M} N={p.N} prec={p.precision} steps={steps} max|mix-alone|={d:.3e}", flush=True)
        assert np.array_equal(phi, aphi) and np.array_equal(con, acon), f"job {jid}: {d:.3e}"
        assert st.energy == e[0] and st.density == e[1]
    srv.close()


def _cli(args, outdir):
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pfc-exp-gui_amd", "pfc-exp-cli")
    r = subprocess.run([cli] + args + ["--outdir", str(outdir)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]


def _same_files(a, b):
    """Same file names; every file byte-identical except the wall-clock
    column of checkpoint_timestamps.txt."""
    na, nb = sorted(os.listdir(a)), sorted(os.listdir(b))
    assert na == nb, (set(na) ^ set(nb))
    for f in na:
        x, y = open(os.path.join(a, f), "rb").read(), open(os.path.join(b, f), "rb").read()
        if f == "checkpoint_timestamps.txt":
            x = [l.split()[:2] for l in x.decode().splitlines()[1:]]
            y = [l.split()[:2] for l in y.decode().splitlines()[1:]]
        assert x == y, f
    return na


def test_served_files_match_cli(hip, tmp_path):
    """A served Misfit job and a served CVD job (FP32, batched beside other
    jobs of their class) write the files pfc-exp-cli writes for the same
    arguments, byte for byte: the kd = 0 fields and a checkpoint every mod
    steps and at the end for Misfit; energy after every step and a
    checkpoint after step t when t % mod == 0 for CVD (cvd.cpp:389-454)."""
    srv = hip.Server(max_slots=3, graph_chunk=5, quantum=10)
    pm = hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4, seed=31)
    pc = hip.default_params(model=hip.CVD, L=64, M=32, precision=hip.FP32, u0=0.05, con0=0.25)
    a = srv.submit(pm, 50, mod=20, outdir=str(tmp_path / "srv_misfit"))
    c = srv.submit(pc, 43, mod=20, outdir=str(tmp_path / "srv_cvd"))
    others = [srv.submit(hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4,
                                            seed=40 + k), 37 + 11 * k) for k in range(2)]
    others.append(srv.submit(hip.default_params(model=hip.CVD, L=64, M=32, precision=hip.FP32, u0=0.07), 29))
    srv.drain(120)
    for jid in [a, c] + others:
        assert srv.poll(jid).state == hip.JOB_DONE, srv.error(jid)
    _cli(["--model", "misfit", "--precision", "fp32", "--L", "64", "--M", "64", "--grainx", "4", "--grainy", "4",
          "--seed", "31", "--steps", "50", "--mod", "20"], tmp_path / "cli_misfit")
    _cli(["--model", "cvd", "--precision", "fp32", "--L", "64", "--M", "32", "--u0", "0.05", "--con0", "0.25",
          "--steps", "43", "--mod", "20"], tmp_path / "cli_cvd")
    names = _same_files(tmp_path / "srv_misfit", tmp_path / "cli_misfit")
    assert {"phi_0_0.vti", "phi_20_0.vti", "phi_40_0.vti", "phi_50_0.vti", "Phimax_50.txt"} <= set(names)
    names = _same_files(tmp_path / "srv_cvd", tmp_path / "cli_cvd")
    assert {"phi_0_0.vti", "phi_20_0.vti", "phi_40_0.vti", "Ql_0.txt", "Ql_20.txt", "Ql_40.txt",
            "hist_40.txt"} <= set(names)
    assert len(open(tmp_path / "srv_cvd" / "energy.txt").read().splitlines()) == 43
    srv.close()


def test_outdir_files_and_cancel(hip, tmp_path):
    srv = hip.Server(max_slots=2, graph_chunk=5, quantum=10)
    p = hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4, seed=31)
    out = str(tmp_path / "job")
    a = srv.submit(p, 50, mod=20, keep_fields=True, outdir=out)
    long_job = srv.submit(p, 10 ** 7)
    time.sleep(0.2)
    srv.cancel(long_job)
    srv.drain(120)
    st = srv.poll(a)
    assert srv.poll(long_job).state == hip.JOB_CANCELLED
    assert st.state == hip.JOB_DONE and st.t_start <= st.t_stepped <= st.t_end
    names = set(os.listdir(out))
    for kd in (0, 20, 40, 50):
        for f in (f"phi_{kd}_0.vti", f"con_{kd}_0.vti", f"phi_{kd}.pvti", f"con_{kd}.pvti"):
            assert f in names, f
    assert {"run_config.txt", "energy.txt", "checkpoint_timestamps.txt", "Phimax_20.txt", "Phimax_50.txt"} <= names
    kds = [int(line.split()[0]) for line in open(os.path.join(out, "energy.txt"))]
    assert kds == [20, 40, 50]
    # the final VTI holds the final fields as float32 (payload after the 4-byte size word)
    phi, _ = srv.result(a)
    raw = open(os.path.join(out, "phi_50_0.vti"), "rb").read()
    pos = 0
    for _ in range(12):  # 12 XML lines, then '_' and the int32 byte count
        pos = raw.index(b"\n", pos) + 1
    assert raw[pos:pos + 1] == b"_"
    start = pos + 1 + 4
    vals = np.frombuffer(raw[start:start + 4 * phi.size], dtype=np.float32).reshape(p.M, p.L)  # [m][l]
    assert np.array_equal(vals, phi.reshape(p.L, p.M).T.astype(np.float32))
    srv.close()


def test_failing_job_fails_alone(hip, tmp_path):
    """Two jobs of one class checkpoint in the same rounds of one batch
    context; the first one's output directory is replaced by a plain file
    right after submission, so its checkpoint files cannot be written.  It
    ends FAILED with a file error; the other ends DONE with its files and the
    same result as run alone (a job's failure is its own, not its batch's)."""
    srv = hip.Server(max_slots=4, graph_chunk=5, quantum=10)
    pa = hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4, seed=31)
    pb = hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4, seed=32)
    bad, good = tmp_path / "bad", tmp_path / "good"
    a = srv.submit(pa, 2000, mod=200, outdir=str(bad))
    b = srv.submit(pb, 2000, mod=200, keep_fields=True, outdir=str(good))
    shutil.rmtree(bad, ignore_errors=True)
    bad.write_text("not a directory\n")
    srv.drain(120)
    sa, sb = srv.poll(a), srv.poll(b)
    print(f"failing job: state {sa.state} error {srv.error(a)!r}; other job: state {sb.state}", flush=True)
    assert sa.state == hip.JOB_FAILED and "files" in srv.error(a)
    assert sb.state == hip.JOB_DONE and sb.steps_done == 2000, srv.error(b)
    phi, con = srv.result(b)
    (aphi, acon), e = _alone(hip, pb, 2000)
    assert np.array_equal(phi, aphi) and np.array_equal(con, acon)
    assert sb.energy == e[0]
    names = set(os.listdir(good))
    assert {"phi_0_0.vti", "phi_1000_0.vti", "phi_2000_0.vti", "Phimax_2000.txt", "energy.txt"} <= names
    srv.close()


def test_request_loop_serving_shapes(hip, oracle):
    """The serving stream's large shapes through the request loop beside
    small jobs: CVD 1024^2 (two jobs in one batch context) and Misfit
    128x128x64, FP32.  Every result equals the same job run alone, bit for
    bit; the short CVD 1024^2 job also matches the FP64 oracle within FP32
    round-off after its 3 steps."""
    srv = hip.Server(max_slots=4, graph_chunk=10, quantum=40)
    jobs = [(hip.default_params(model=hip.CVD, L=1024, M=1024, precision=hip.FP32, u0=0.05, con0=0.25), 3),
            (hip.default_params(model=hip.CVD, L=1024, M=1024, precision=hip.FP32, u0=0.06, con0=0.3), 57),
            (hip.default_params(model=hip.MISFIT, L=128, M=128, N=64, precision=hip.FP32, grainx=8, grainy=8,
                                seed=77), 31),
            (hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4, seed=5), 90),
            (hip.default_params(model=hip.CVD, L=128, M=128, precision=hip.FP32), 45)]
    ids = []
    for k, (p, steps) in enumerate(jobs):
        ids.append(srv.submit(p, steps, keep_fields=True))
        if k == 2:
            time.sleep(0.05)
    srv.drain(300)
    for jid, (p, steps) in zip(ids, jobs):
        st = srv.poll(jid)
        assert st.state == hip.JOB_DONE, srv.error(jid)
        phi, con = srv.result(jid)
        (aphi, acon), e = _alone(hip, p, steps)
        print(f"job {jid} {p.L}x{p.M}x{p.N} model={p.model} steps={steps} create_s={st.create_s:.3f} "
              f"queue={st.t_start - st.t_submit:.3f} stepping={st.t_stepped - st.t_start:.3f}", flush=True)
        assert np.array_equal(phi, aphi) and np.array_equal(con, acon), jid
        assert st.energy == e[0] and st.density == e[1]
    p, steps = jobs[0]
    phi, con = srv.result(ids[0])
    m = oracle.model(1, 1024, 1024, 1, u0=0.05, con0=0.25)
    osim = oracle.Sim(m)
    osim.run(steps)
    ophi, ocon = osim.fields()
    err = max(np.abs(phi - ophi).max(), np.abs(con - ocon).max())
    print(f"served CVD 1024^2 FP32 vs oracle after {steps} steps: {err:.3e}", flush=True)
    assert err < 1e-5
    srv.close()


def test_cancel_after_done_keeps_result(hip):
    """A cancellation that arrives after a job has completed leaves it DONE,
    with its steps and final fields (those of the same job run alone)."""
    srv = hip.Server(max_slots=2, graph_chunk=5, quantum=10)
    p = hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4, seed=41)
    a = srv.submit(p, 30, keep_fields=True)
    srv.drain(120)
    assert srv.poll(a).state == hip.JOB_DONE
    phi0, con0 = srv.result(a)
    srv.cancel(a)
    srv.drain(120)
    st = srv.poll(a)
    assert st.state == hip.JOB_DONE and st.steps_done == 30
    phi1, con1 = srv.result(a)
    assert np.array_equal(phi0, phi1) and np.array_equal(con0, con1)
    (ref_phi, ref_con), _ = _alone(hip, p, 30)
    assert np.array_equal(phi1, ref_phi) and np.array_equal(con1, ref_con)
    srv.close()
