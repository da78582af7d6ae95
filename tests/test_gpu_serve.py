"""Request loop with continuous batching on the GPU.

Jobs that arrive while others are stepping are admitted into free slots of
their class's batch context (idle slots masked); each result is the same as
the job run alone in a one-instance context of the same precision -- bit
for bit, since every instance runs the same kernels on its own data.
"""
import os
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


def test_outdir_files_and_cancel(hip, tmp_path):
    srv = hip.Server(max_slots=2, graph_chunk=5, quantum=10)
    p = hip.default_params(model=hip.MISFIT, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4, seed=31)
    out = str(tmp_path / "job")
    a = srv.submit(p, 50, mod=20, keep_fields=True, outdir=out)
    long_job = srv.submit(p, 10 ** 7)
    time.sleep(0.2)
    srv.cancel(long_job)
    q = hip.default_params(model=hip.CVD, L=32, M=32, precision=hip.FP32)
    c = srv.submit(q, 12, mod=5, outdir=str(tmp_path / "cvd"))
    srv.drain(120)
    assert srv.poll(long_job).state == hip.JOB_CANCELLED
    assert srv.poll(a).state == hip.JOB_DONE and srv.poll(c).state == hip.JOB_DONE
    names = set(os.listdir(out))
    for kd in (20, 40, 50):
        for f in (f"phi_{kd}_0.vti", f"con_{kd}_0.vti", f"phi_{kd}.pvti", f"con_{kd}.pvti", f"Phimax_{kd}.txt"):
            assert f in names, f
    assert {"run_config.txt", "energy.txt", "checkpoint_timestamps.txt"} <= names
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
    cnames = set(os.listdir(tmp_path / "cvd"))
    assert {"phimax.txt", "phiball_5.cfg", "Phimax_12.txt", "phi_10_0.vti", "hist_5.txt", "Grq6_12.txt"} <= cnames
    assert not any(n.startswith("Ql_") for n in cnames)  # Q6_GLO only when kd % 20 == 0
    assert len(open(tmp_path / "cvd" / "hist_10.txt").read().splitlines()) == int(np.float32(32) * np.float32(q.dx) / np.float32(0.04))
    srv.close()
