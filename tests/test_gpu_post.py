"""GPU Phimax / strain parity: the extraction kernels against the oracle on
the same fields.  The GPU Phimax reads the device field; the oracle gets the
same values through pfc_get_fields, so every atom must match bit for bit
(integer/index work and IEEE double/float arithmetic without contraction)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(hip, kind, prec, L=64, M=64, steps=40, **kw):
    p = hip.default_params(model=kind, L=L, M=M, precision=prec, **kw)
    phi, con = hip.init_fields(p)
    sim = hip.Sim(p)
    sim.set_fields(phi, con)
    sim.step(steps)
    return p, sim


@pytest.mark.parametrize("prec", [0, 1])
def test_phimax_misfit(hip, oracle, prec):
    p, sim = _state(hip, 0, prec, grainx=4, grainy=4)
    g = sim.phimax_misfit()
    phi, con = sim.get_fields()
    o = oracle.phimax_misfit(p.L, p.M, 1, p.dx, p.con0, phi, con)
    print(f"prec={prec} misfit atoms gpu={len(g[0])} oracle={len(o[0])}", flush=True)
    assert len(g[0]) == len(o[0]) > 0
    for a, b in zip(g, o):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [0, 1])
def test_phimax_cvd(hip, oracle, prec):
    p, sim = _state(hip, 1, prec)
    px, py, pz, pp, phic, pmax = sim.phimax_cvd()
    phi, _ = sim.get_fields()
    o = oracle.phimax_cvd(p.L, p.M, 1, p.dx, p.u0, phi)
    print(f"prec={prec} cvd atoms gpu={len(px)} oracle={len(o[0])}", flush=True)
    assert phic == o[4] and pmax == o[5]
    assert len(px) == len(o[0]) > 0
    for a, b in zip((px, py, pz, pp), o[:4]):
        assert np.array_equal(a, b)


def test_phimax_elastic_and_strain(hip, oracle):
    p, sim = _state(hip, 0, 0, L=128, M=128, steps=60, grainx=4, grainy=4)
    phi, _ = sim.get_fields()
    f = phi.astype(np.float32)
    g = hip.phimax_elastic(p.L, p.M, 1, p.dx, f)
    o = oracle.phimax_elastic(p.L, p.M, 1, p.dx, f)
    print(f"elastic atoms gpu={len(g[0])} oracle={len(o[0])}", flush=True)
    assert len(g[0]) == len(o[0]) > 0
    for a, b in zip(g, o):
        assert np.array_equal(a, b)
    s, nb = hip.strain(p.L, p.M, 1, p.dx, 1.7, g[0], g[1], g[2])
    os_, onb = oracle.strain(p.L, p.M, 1, p.dx, 1.7, o[0], o[1], o[2])
    assert np.array_equal(nb, onb)
    assert np.array_equal(s, os_)


def test_elastic_3d_field(hip, oracle):
    """N > 1 exercises the z stencil and its wrap arithmetic."""
    rng = np.random.default_rng(7)
    L, M, N = 16, 16, 8
    x = np.arange(L)[:, None, None]
    y = np.arange(M)[None, :, None]
    z = np.arange(N)[None, None, :]
    f = (np.cos(2 * np.pi * x / 4) * np.cos(2 * np.pi * y / 4) * np.cos(2 * np.pi * z / 4)
         + 0.05 * rng.standard_normal((L, M, N))).astype(np.float32).ravel()
    g = hip.phimax_elastic(L, M, N, 0.125, f)
    o = oracle.phimax_elastic(L, M, N, 0.125, f)
    assert len(g[0]) == len(o[0]) > 0
    for a, b in zip(g, o):
        assert np.array_equal(a, b)
