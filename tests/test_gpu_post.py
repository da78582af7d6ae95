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


@pytest.mark.parametrize("prec,L,g0", [(0, 64, 4), (1, 64, 4), (0, 512, 16), (1, 512, 16)])
def test_phimax_misfit(hip, oracle, prec, L, g0):
    """Atom extraction bit-identical to the oracle on the same fields, also at
    the benchmark grid (512^2, 16x16 grains)."""
    p, sim = _state(hip, 0, prec, L=L, M=L, grainx=g0, grainy=g0)
    g = sim.phimax_misfit()
    phi, con = sim.get_fields()
    o = oracle.phimax_misfit(p.L, p.M, 1, p.dx, p.con0, phi, con)
    print(f"prec={prec} misfit atoms gpu={len(g[0])} oracle={len(o[0])}", flush=True)
    assert len(g[0]) == len(o[0]) > 0
    for a, b in zip(g, o):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec,L", [(0, 64), (1, 64), (0, 1024), (1, 1024)])
def test_phimax_cvd(hip, oracle, prec, L):
    """CVD atoms and the phic/pmax pair bit-identical to the oracle on the
    same field, also at the checkpoint run's grid (1024^2)."""
    p, sim = _state(hip, 1, prec, L=L, M=L)
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


def test_cryst_order(hip, oracle):
    """q4/q6 of the atoms of a grown crystal: neighbour lists and counts exact,
    q within float transcendental round-off (device vs host cosf/sinf)."""
    p, sim = _state(hip, 0, 0, L=128, M=128, steps=60, grainx=4, grainy=4)
    phi, _ = sim.get_fields()
    g = hip.phimax_elastic(p.L, p.M, 1, p.dx, phi.astype(np.float32))
    q4, q6, nb = hip.cryst_order(p.L, p.M, 1, p.dx, g[0], g[1], g[2])
    o4, o6, onb = oracle.cryst_order(p.L, p.M, 1, p.dx, g[0], g[1], g[2])
    d = max(np.abs(q4 - o4).max(), np.abs(q6 - o6).max())
    print(f"cryst_order atoms={len(q4)} max|dq|={d:.3e} nb range {nb.min()}..{nb.max()}", flush=True)
    assert len(q4) > 100
    assert np.array_equal(nb, onb)
    assert d < 1e-5
    assert np.allclose(q6, np.sqrt(2.0) * q4, rtol=1e-6)


def test_cryst_order_3d_lattice(hip, oracle):
    """A periodic fcc lattice (nearest distance 1, 12 neighbours) in a 3-D
    box, jittered: every atom ordered, GPU = oracle."""
    rng = np.random.default_rng(3)
    a, cells, dx = np.sqrt(2.0), 6, np.float32(np.sqrt(2.0) / 8)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = np.array([(np.array(c) + b) * a for c in np.ndindex(cells, cells, cells) for b in base])
    pts += 0.01 * rng.standard_normal(pts.shape)
    L = M = N = cells * 8
    q4, q6, nb = hip.cryst_order(L, M, N, float(dx), pts[:, 0], pts[:, 1], pts[:, 2])
    o4, o6, onb = oracle.cryst_order(L, M, N, float(dx), pts[:, 0], pts[:, 1], pts[:, 2])
    assert np.array_equal(nb, onb) and (nb == 12).all()
    assert max(np.abs(q4 - o4).max(), np.abs(q6 - o6).max()) < 1e-5
    assert np.all(q4 > 0.5)


def _q6_close(g, o, tol=1e-5):
    nb, ql, qg, hq = g
    onb, oql, oqg, ohq = o
    assert np.array_equal(nb, onb)
    nan = np.isnan(oql)  # an atom without neighbours has no local Q6 (0/0)
    assert np.array_equal(np.isnan(ql), nan)
    d = np.abs(ql[~nan] - oql[~nan])
    assert (d.max() if d.size else 0.0) < tol and abs(qg - oqg) < tol
    # a local value within tol of a bin edge (multiples of 0.05) may change bin
    edge = np.abs(oql / 0.05 - np.round(oql / 0.05)) * 0.05 < tol
    assert np.abs(hq.astype(np.int64) - ohq.astype(np.int64)).sum() <= 2 * edge.sum()


def test_cvd_gr_q6_field(hip, oracle):
    """GR and Q6_GLO of the atoms of a grown 3-D CVD field: histogram exact,
    neighbour counts exact, Q6 within float transcendental round-off."""
    p = hip.default_params(model=1, L=32, M=32, N=16, precision=hip.FP64)
    phi, con = hip.init_fields(p)
    sim = hip.Sim(p)
    sim.set_fields(phi, con)
    sim.step(20)
    px, py, pz, pp, phic, pmax = sim.phimax_cvd()
    print(f"cvd atoms={len(px)}", flush=True)
    h = hip.cvd_gr(p.L, p.M, p.N, p.dx, px, py, pz)
    oh = oracle.cvd_gr(p.L, p.M, p.N, p.dx, px, py, pz)
    assert len(h) == len(oh) and h.sum() > 0
    assert np.array_equal(h, oh)
    _q6_close(hip.cvd_q6(p.L, p.M, p.N, p.dx, px, py, pz), oracle.cvd_q6(p.L, p.M, p.N, p.dx, px, py, pz))


@pytest.mark.parametrize("L,N,n", [(64, 64, 4000), (4096, 16, 3000)])
def test_cvd_gr_q6_sizes(hip, oracle, L, N, n):
    """Many i/j tiles (the triangle's block skip), and a histogram too large
    for LDS (L*dx/delr > 12288 bins: global atomics)."""
    rng = np.random.default_rng(L)
    M, dx = 64, 0.125
    f = rng.random((n, 3)).astype(np.float32)
    f[:, 0] *= 64.0 / L                   # keep the atoms dense along x
    h = hip.cvd_gr(L, M, N, dx, f[:, 0], f[:, 1], f[:, 2])
    oh = oracle.cvd_gr(L, M, N, dx, f[:, 0], f[:, 1], f[:, 2])
    print(f"L={L} n={n} bins={len(h)} pairs={h.sum() // 2}", flush=True)
    assert np.array_equal(h, oh)
    _q6_close(hip.cvd_q6(L, M, N, dx, f[:, 0], f[:, 1], f[:, 2]), oracle.cvd_q6(L, M, N, dx, f[:, 0], f[:, 1], f[:, 2]))


def test_cvd_gr_q6_checkpoint_grid(hip, oracle):
    """GR and Q6_GLO at the checkpoint run's size: the atoms of a CVD 1024^2
    FP32 field after 500 steps (the first checkpoint of that run)."""
    p = hip.default_params(model=1, L=1024, M=1024, precision=hip.FP32, u0=0.05, con0=0.25)
    phi, con = hip.init_fields(p)
    sim = hip.Sim(p)
    sim.set_fields(phi, con)
    sim.step(500)
    px, py, pz, pp, phic, pmax = sim.phimax_cvd()
    sim.close()
    print(f"cvd 1024^2 atoms={len(px)}", flush=True)
    assert len(px) > 1000
    h = hip.cvd_gr(p.L, p.M, p.N, p.dx, px, py, pz)
    oh = oracle.cvd_gr(p.L, p.M, p.N, p.dx, px, py, pz)
    assert np.array_equal(h, oh) and h.sum() > 0
    _q6_close(hip.cvd_q6(p.L, p.M, p.N, p.dx, px, py, pz), oracle.cvd_q6(p.L, p.M, p.N, p.dx, px, py, pz))


def test_cvd_q6_lattice(hip, oracle):
    """Jittered fcc (12 neighbours in the window): GPU = oracle, local and
    global Q6 near the fcc value 0.5745."""
    rng = np.random.default_rng(4)
    a, cells, L = 0.9 * 2 ** 0.5, 6, 48
    dx = float(np.float32(cells * a / L))
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = np.array([(np.array(c) + b) * a for c in np.ndindex(cells, cells, cells) for b in base])
    pts += 0.01 * rng.standard_normal(pts.shape) + 0.05
    f = (pts / np.float32(L * np.float32(dx))).astype(np.float32)
    g = hip.cvd_q6(L, L, L, dx, f[:, 0], f[:, 1], f[:, 2])
    o = oracle.cvd_q6(L, L, L, dx, f[:, 0], f[:, 1], f[:, 2])
    _q6_close(g, o)
    assert (g[0] == 12).all() and abs(g[2] - 0.5745) < 0.02


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
