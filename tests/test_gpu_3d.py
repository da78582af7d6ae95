"""3-D grids (N > 1) and the per-instance step mask on the GPU.

3-D steps run the depth pass (FFT along n, with the packed k_m = 0 / M/2
column split along n) on both sides of the column pass; parity is against
the oracle's full 3-D FFT restatement (pfo_fft3d) on the reference's own
initial fields, with the 2-D tolerances: FP64 1e-11, FP32 2e-4, noise FP64
1e-7.  The CVD maxima extraction on a 3-D field is bit-exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair3(hip, oracle, kind, L, M, N, prec, narrow=0, **kw):
    p = hip.default_params(model=kind, L=L, M=M, N=N, precision=prec, **kw)
    p.narrow = narrow
    m = oracle.model(kind, L, M, N, narrow=narrow, **kw)
    phi, con = hip.init_fields(p)
    ophi, ocon = oracle.init_fields(m)
    assert np.array_equal(phi, ophi) and np.array_equal(con, ocon), "initial fields differ"
    sim = hip.Sim(p)
    sim.set_fields(phi, con)
    osim = oracle.Sim(m, phi, con)
    return p, sim, osim


def _err(sim, osim):
    phi, con = sim.get_fields()
    ophi, ocon = osim.fields()
    return max(np.abs(phi - ophi).max(), np.abs(con - ocon).max())


@pytest.mark.parametrize("kind,L,M,N,kw", [
    (0, 32, 32, 16, dict(grainx=2, grainy=2)),
    (1, 64, 32, 32, {}),
    (0, 32, 64, 64, dict(grainx=2, grainy=4)),
    (1, 32, 32, 128, {}),
])
def test_3d_fp64_exact(hip, oracle, kind, L, M, N, kw):
    _, sim, osim = _pair3(hip, oracle, kind, L, M, N, hip.FP64, **kw)
    for n in (1, 7):
        sim.step(n)
        osim.run(n)
        err = _err(sim, osim)
        print(f"kind={kind} {L}x{M}x{N} t={osim.t} fp64 err={err:.3e}", flush=True)
        assert err < 1e-11


@pytest.mark.parametrize("kind", [0, 1])
def test_3d_fp32(hip, oracle, kind):
    kw = dict(grainx=2, grainy=2) if kind == 0 else {}
    _, sim, osim = _pair3(hip, oracle, kind, 32, 64, 16, hip.FP32, narrow=1, **kw)
    sim.step(30)
    osim.run(30)
    err = _err(sim, osim)
    print(f"kind={kind} 3-D fp32 err={err:.3e}", flush=True)
    assert err < 2e-4


@pytest.mark.parametrize("kind", [0, 1])
def test_3d_serving_grid(hip, oracle, kind):
    """The request loop's 3-D grid, 128 x 128 x 64, FP32, against the oracle
    (a few steps: the oracle takes about a second per step here)."""
    kw = dict(grainx=2, grainy=2) if kind == 0 else {}
    _, sim, osim = _pair3(hip, oracle, kind, 128, 128, 64, hip.FP32, narrow=1, **kw)
    sim.step(3)
    osim.run(3)
    err = _err(sim, osim)
    print(f"kind={kind} 128x128x64 fp32 err={err:.3e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("prec", [0, 1])
def test_3d_energy(hip, oracle, prec):
    _, sim, osim = _pair3(hip, oracle, 0, 32, 32, 16, prec, narrow=1, grainx=2, grainy=2)
    sim.step(10)
    osim.run(10)
    before = sim.get_fields()
    e = sim.energy()[0]
    oe, orho = osim.energy()
    print(f"prec={prec} 3-D energy {e} oracle {oe} {orho}", flush=True)
    assert abs(e[0] - oe) < 1e-6 * max(1.0, abs(oe)) + (1e-9 if prec == 0 else 1e-6)
    assert abs(e[1] - orho) < 1e-9
    after = sim.get_fields()  # the energy evaluation leaves the state untouched
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])


def test_3d_step_energy_trace(hip, oracle):
    _, sim, osim = _pair3(hip, oracle, 1, 32, 32, 16, hip.FP64, narrow=1)
    tr = sim.step_energy(4)
    for t in range(4):
        osim.run(1)
        oe, orho = osim.energy()
        assert abs(tr[t, 0, 0] - oe) < 1e-9 + 1e-6 * abs(oe), (t, tr[t, 0], oe)
        assert abs(tr[t, 0, 1] - orho) < 1e-9


def test_3d_noise(hip, oracle):
    _, sim, osim = _pair3(hip, oracle, 0, 32, 32, 16, hip.FP64, grainx=2, grainy=2, axx=0.001)
    sim.step(4)
    sim.sync()
    osim.run(4)
    err = _err(sim, osim)
    print(f"3-D noise fp64 err={err:.3e}", flush=True)
    assert err < 1e-7


@pytest.mark.parametrize("prec", [0, 1])
def test_3d_phimax_cvd(hip, oracle, prec):
    p, sim, _ = _pair3(hip, oracle, 1, 32, 32, 16, prec)
    sim.step(20)
    px, py, pz, pp, phic, pmax = sim.phimax_cvd()
    phi, _ = sim.get_fields()
    o = oracle.phimax_cvd(p.L, p.M, p.N, p.dx, p.u0, phi)
    print(f"prec={prec} 3-D cvd atoms gpu={len(px)} oracle={len(o[0])}", flush=True)
    assert phic == o[4] and pmax == o[5]
    assert len(px) == len(o[0]) > 0
    for a, b in zip((px, py, pz, pp), o[:4]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("N", [1, 16])
def test_step_mask(hip, N):
    """An idle instance keeps its state and its noise step index; the others
    advance exactly as when run alone."""
    p = hip.default_params(model=0, L=32, M=64, N=N, precision=hip.FP32, grainx=2, grainy=2, axx=0.001)
    fields = []
    for s in range(3):
        p.seed = 77 + s
        fields.append(hip.init_fields(p))
    b = hip.Sim(p, batch=3)
    for i, (phi, con) in enumerate(fields):
        b.set_fields(phi, con, inst=i)
        b.set_seed(77 + i, inst=i)
    b.set_graph(5)
    b.step(10)
    b.set_active(1, False)
    b.step(10)
    mid = b.get_fields(1)
    b.set_active(1, True)
    b.set_active(2, False)
    b.step(10)
    b.sync()
    want_steps = (30, 20, 20)
    for i, (phi, con) in enumerate(fields):
        p.seed = 77 + i
        s1 = hip.Sim(p)
        s1.set_fields(phi, con)
        s1.set_seed(77 + i)
        s1.step(want_steps[i] if i != 1 else 10)
        if i == 1:
            x = s1.get_fields()
            assert np.array_equal(x[0], mid[0]) and np.array_equal(x[1], mid[1])
            s1.step(10)
        x = s1.get_fields()
        y = b.get_fields(i)
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]), f"instance {i}"


def test_concurrent_contexts_bit_identical(hip):
    """Contexts stepping at the same time on their own streams (a 3-D FP64
    batch beside an FP32 batch with the noise term) give the results each
    gives alone: nothing is shared between contexts."""
    pa = hip.default_params(model=0, L=32, M=32, N=16, precision=hip.FP64, grainx=2, grainy=2, seed=9)
    pb = hip.default_params(model=0, L=64, M=64, precision=hip.FP32, grainx=4, grainy=4, seed=77, axx=0.001)

    def make(p, B):
        s = hip.Sim(p, batch=B)
        phi, con = hip.init_fields(p)
        for i in range(B):
            s.set_fields(phi, con, inst=i)
        s.set_seed(p.seed)
        return s

    ref = {}
    for key, p in (("a", pa), ("b", pb)):
        s = make(p, 1)
        s.step(30)
        ref[key] = s.get_fields()
        s.close()
    for graph in (8, 0):
        for rep in range(3):
            a, b = make(pa, 4), make(pb, 4)
            for i in range(1, 4):
                a.set_active(i, False)
                b.set_active(i, False)
            if graph:
                a.set_graph(graph)
                b.set_graph(graph)
            for _ in range(30 // 6):
                a.step(6)
                b.step(6)
            a.sync()
            b.sync()
            for key, s in (("a", a), ("b", b)):
                x = s.get_fields(0)
                assert np.array_equal(x[0], ref[key][0]) and np.array_equal(x[1], ref[key][1]), (key, graph, rep)
            a.close()
            b.close()
