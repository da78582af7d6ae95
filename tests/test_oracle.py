"""Pins for the CPU oracle (the parity checker).

Each check is chosen so that a plausible mistake (a dropped term, a wrong sign
or index, a transposed operand, a missing normalisation) fails: published
known-answer values, numpy's pocketfft, derivative identities of the free
energy, exact invariants of the scheme, planted maxima with known positions,
and small cases with hand-computable answers.  The reference has no tests or
fixtures of its own for this path.
"""
import ctypes
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --------------------------------------------------------------------------
# RNG and initial fields
# --------------------------------------------------------------------------
def test_splitmix64_kat(oracle):
    lib = oracle.load()
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    st = ctypes.c_uint64(0)
    got = [lib.pfo_splitmix64(ctypes.byref(st)) for _ in range(3)]
    assert got == [int(h, 16) for h in kat["splitmix64_state0"]]


def test_u01_is_53_bit(oracle):
    lib = oracle.load()
    st = ctypes.c_uint64(12345)
    st2 = ctypes.c_uint64(12345)
    u = lib.pfo_rng_u01(ctypes.byref(st))
    raw = lib.pfo_splitmix64(ctypes.byref(st2))
    assert u == (raw >> 11) * 2.0 ** -53
    assert 0.0 <= u < 1.0


def test_gauss_noise_stream(oracle):
    """The --axx noise stream (Marsaglia polar over splitmix64 uniforms,
    misfit.cpp:951-981), pinned by what the method fixes: N(0,1) moments
    (standard errors 1e-3 for the mean, 1.4e-3 for the variance, 1e-2 for the
    4th moment at 2^20 values), the Kolmogorov distance to the normal CDF, and
    an acceptance rate of pi/4 (the unit disc in the square).  A dropped -2 or
    /r in the radius, or a Box-Muller radius, changes the variance."""
    from scipy.stats import kstest
    n = 1 << 20
    x, pairs = oracle.gauss(20240611, 3, n)
    assert np.all(np.isfinite(x))
    assert abs(x.mean()) < 5e-3
    assert abs(x.var() - 1.0) < 7e-3
    assert abs((x ** 4).mean() - 3.0) < 5e-2
    assert kstest(x, "norm").statistic < 3e-3
    assert abs(n / pairs - np.pi / 4) < 3e-3
    y, _ = oracle.gauss(20240611, 3, 1000)
    assert np.array_equal(x[:1000], y)  # a step's stream starts afresh from (seed, timestep)
    for other in (oracle.gauss(20240611, 4, 1000)[0], oracle.gauss(7, 3, 1000)[0]):
        assert np.abs(other - y).max() > 1.0


def test_grain_angles_kat(oracle):
    lib = oracle.load()
    kat = json.load(open(os.path.join(GOLD, "kat.json")))["grain_theta_deg"]
    m = oracle.model(0, kat["L"], kat["M"], 1, grainx=kat["grainx"], grainy=kat["grainy"], seed=kat["seed"])
    for i, j, deg in kat["cases"]:
        th = lib.pfo_grain_theta(ctypes.byref(m), i, j, 0)
        # theta is stored as float; 3.1415926 is the reference's pi
        assert abs(th * 180 / 3.1415926 - deg) < 2e-6 * abs(deg)


def test_misfit_init_structure(oracle):
    m = oracle.model(0, 256, 256, 1)
    phi, con = oracle.init_fields(m)
    phi = phi.reshape(256, 256)
    assert np.all(con == np.float32(0.2))
    # grains only in rows m < 160 (misfit.cpp:616): the rest stays at u0
    assert np.all(phi[:, 160:] == np.float32(0.05))
    assert np.any(phi[:, :160] != np.float32(0.05))
    # 13x13-point patches: half width 256*0.125/16/2 - 0.2 = 0.8 -> |d| < 6.4 pts
    # around centres dn*i + dn/2 = 16 i + 8
    col = phi[:, 8]
    seeded = np.nonzero(col != np.float32(0.05))[0]
    assert set(seeded) <= set(k for c in range(8, 256, 16) for k in range(c - 6, c + 7))


def test_cvd_init_two_phases(oracle):
    m = oracle.model(1, 64, 64, 1)
    phi, con = oracle.init_fields(m)
    con = con.reshape(64, 64)
    assert np.all(con[:, :32] == np.float32(0.2))
    assert np.all(con[:, 32:] == np.float32(1) - np.float32(0.2))
    phi = phi.reshape(64, 64)
    # square lattice of period a=1 -> 8 points at dx=0.125 (lower half)
    assert np.allclose(phi[8:, :32], phi[:-8, :32], atol=1e-5)


# --------------------------------------------------------------------------
# FFT against numpy pocketfft golden vectors
# --------------------------------------------------------------------------
def test_fft_1d_golden(oracle):
    g = np.load(os.path.join(GOLD, "fft_golden.npz"))
    lib = oracle.load()
    for key in g.files:
        if not key.startswith("in1d_"):
            continue
        n = int(key.split("_")[1])
        for sign, ref in ((-1, g[f"fwd1d_{n}"]), (1, g[f"bwd1d_{n}"])):
            a = g[key].astype(np.complex128).copy()
            lib.pfo_fft1d(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, sign)
            assert np.abs(a - ref).max() < 1e-12 * np.abs(ref).max(), (n, sign)


def test_fft_3d_golden(oracle):
    g = np.load(os.path.join(GOLD, "fft_golden.npz"))
    for key in g.files:
        if not key.startswith("in3d_"):
            continue
        shape = tuple(int(s) for s in key.split("_")[1].split("x"))
        out = oracle.fft3d(g[key].ravel(), *shape, -1)
        ref = g["fwd3d_" + key.split("_")[1]].ravel()
        assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max(), shape


def test_fft_roundtrip_and_impulse(oracle):
    x = np.zeros(96, np.complex128)
    x[0] = 1
    y = oracle.fft3d(x, 96, 1, 1, -1)
    assert np.allclose(y, 1.0, atol=1e-15)
    rng = np.random.default_rng(1)
    z = rng.standard_normal(384) + 1j * rng.standard_normal(384)
    back = oracle.fft3d(oracle.fft3d(z, 384, 1, 1, -1), 384, 1, 1, 1) / 384
    assert np.abs(back - z).max() < 1e-13


# --------------------------------------------------------------------------
# k tables and C2 kernels
# --------------------------------------------------------------------------
def _ktabs(oracle, m):
    lib = oracle.load()
    kl = np.zeros(m.L, np.float32)
    km = np.zeros(m.M, np.float32)
    kn = np.zeros(m.N, np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.pfo_ktables(ctypes.byref(m), kl.ctypes.data_as(fp), km.ctypes.data_as(fp), kn.ctypes.data_as(fp))
    return kl, km, kn


def test_ktables(oracle):
    m = oracle.model(0, 384, 64, 1)
    kl, km, kn = _ktabs(oracle, m)
    L = 384
    base = 2.0 * 3.1415926 / np.float64(np.float32(L) * np.float32(0.125))
    assert kl[1] == np.float32(base)
    assert kl[L // 2] == np.float32((L // 2) * base)          # Nyquist is positive
    assert np.array_equal(kl[L // 2 + 1:], -kl[1:L // 2][::-1])  # exact odd symmetry
    assert kn[0] == 0.0


def test_c2_kernel_properties(oracle):
    lib = oracle.load()
    m = oracle.model(0, 64, 64, 1)
    c2a = np.zeros(64 * 64, np.float32)
    c2b = np.zeros(64 * 64, np.float32)
    kc1 = ctypes.c_float()
    kc2 = ctypes.c_float()
    fp = ctypes.POINTER(ctypes.c_float)
    lib.pfo_c2(ctypes.byref(m), c2a.ctypes.data_as(fp), c2b.ctypes.data_as(fp), ctypes.byref(kc1), ctypes.byref(kc2))
    # kc is where the two Gaussian branches cross: both branches equal there
    for kc, k1, k2, g1, g2 in ((kc1.value, m.ka1, m.ka2, m.gaa1, m.gaa2), (kc2.value, m.kb1, m.kb2, m.gab1, m.gab2)):
        assert k1 < kc < k2
        b1 = np.exp(-(kc - k1) ** 2 / (2 * g1 * g1))
        b2 = np.exp(-(kc - k2) ** 2 / (2 * g2 * g2))
        assert abs(b1 - b2) < 1e-5
    kl, km, _ = _ktabs(oracle, m)
    k = np.sqrt(kl[:, None].astype(np.float64) ** 2 + km[None, :] ** 2)
    a = c2a.reshape(64, 64)
    # peak value exp(-(sig/0.55)^2) at |k| = ka1, radial symmetry, C2 <= peak
    peak = np.exp(-(0.05 / 0.55) ** 2)
    assert a.max() <= peak * (1 + 1e-6)
    i = np.unravel_index(np.argmin(np.abs(k - m.ka1)), k.shape)
    assert abs(a[i] - peak * np.exp(-(k[i] - m.ka1) ** 2 / (2 * 0.8 ** 2))) < 1e-6
    assert np.array_equal(a[1:, :], a[1:, :][::-1, :][::-1, :])


# --------------------------------------------------------------------------
# Pointwise terms: derivative identities of the free energy
#   F(phi, c) = -nu phi^3/6 + xi phi^4/12 + (phi+1) f(c) + phi/2 * in3(c)
#   in1 + in2 = dF/dphi restricted to the local terms, inc = dF/dc
# --------------------------------------------------------------------------
def test_nonlinear_terms_are_energy_derivatives(oracle):
    lib = oracle.load()
    m = oracle.model(0, 8, 8, 1, narrow=0)
    mp = ctypes.byref(m)
    nu, xi, w, c0 = float(np.float32(1.4)), 1.0, float(np.float32(0.02)), 0.5

    def fmix(c):
        return w * (c * np.log(c / c0) + (1 - c) * np.log((1 - c) / (1 - c0)))

    rng = np.random.default_rng(3)
    for _ in range(50):
        ph, c = rng.uniform(-0.6, 1.1), rng.uniform(0.05, 0.95)
        a, b = rng.uniform(-0.3, 0.3, 2)
        h = 1e-6

        def F(p_, c_):
            return -nu * p_ ** 3 / 6 + xi * p_ ** 4 / 12 + (p_ + 1) * fmix(c_) + 0.5 * p_ * lib.pfo_in3(c_, a, b)

        dphi = (F(ph + h, c) - F(ph - h, c)) / (2 * h) - 0.5 * lib.pfo_in3(c, a, b)
        dc = (F(ph, c + h) - F(ph, c - h)) / (2 * h)
        assert abs(lib.pfo_in1(mp, ph) + lib.pfo_in2(mp, c) - dphi) < 1e-7
        assert abs(lib.pfo_inc(mp, ph, c, a, b) - dc) < 1e-7


def test_inc_narrowing_is_float(oracle):
    lib = oracle.load()
    m1 = oracle.model(0, 8, 8, 1, narrow=1)
    m0 = oracle.model(0, 8, 8, 1, narrow=0)
    rng = np.random.default_rng(4)
    for _ in range(50):
        ph, c, a, b = rng.uniform(-0.6, 1.1), rng.uniform(0.05, 0.95), *rng.uniform(-0.3, 0.3, 2)
        v1 = lib.pfo_inc(ctypes.byref(m1), ph, c, a, b)
        v0 = lib.pfo_inc(ctypes.byref(m0), ph, c, a, b)
        assert v1 == float(np.float32(v1))
        assert abs(v1 - v0) <= 4 * np.spacing(np.float32(abs(v0) + 0.1))


def test_update_closed_form(oracle):
    lib = oracle.load()
    dp = ctypes.POINTER(ctypes.c_double)
    for kind, mn, mc in ((0, 0.1, 10.0), (1, 1.0, 1.0)):
        m = oracle.model(kind, 8, 8, 1, narrow=0)
        for ks in (0.0, 1.0, 39.47, 600.0):
            z = np.array([3.0, -2.0])
            o = np.array([0.5, 0.25])
            lib.pfo_update_phi(ctypes.byref(m), ctypes.c_float(ks), z.ctypes.data_as(dp), o.ctypes.data_as(dp))
            mk = np.float32(np.float32(mn) * np.float32(0.05)) * np.float32(ks)
            exp = (np.array([3.0, -2.0]) - np.float64(mk) * o) / (1 + np.float64(mk))
            assert np.allclose(z, exp, rtol=1e-15)
            zc = np.array([3.0, -2.0])
            lib.pfo_update_con(ctypes.byref(m), ctypes.c_float(ks), zc.ctypes.data_as(dp), o.ctypes.data_as(dp))
            den = 1 + 0.5 * mc * 0.05 * ks * ks
            expc = (np.array([3.0, -2.0]) - mc * 0.05 * ks * o) / den
            assert np.allclose(zc, expc, rtol=1e-6)
            # pure relaxation (no nonlinear drive) never amplifies a mode
            zz = np.array([3.0, -2.0])
            lib.pfo_update_con(ctypes.byref(m), ctypes.c_float(ks), zz.ctypes.data_as(dp),
                               np.zeros(2).ctypes.data_as(dp))
            assert np.all(np.abs(zz) <= np.array([3.0, 2.0]))


# --------------------------------------------------------------------------
# Whole step: invariants and an independent numpy restatement
# --------------------------------------------------------------------------
@pytest.mark.parametrize("kind", [0, 1])
def test_mass_conserved_and_energy_decreases(oracle, kind):
    kw = dict(grainx=4, grainy=4) if kind == 0 else {}
    m = oracle.model(kind, 64, 64, 1, **kw)
    s = oracle.Sim(m)
    s.run(1)
    p1, c1 = s.fields()
    e_prev = s.energy()[0]
    for _ in range(4):
        s.run(10)
        p, c = s.fields()
        # the rounded k=0 coefficient is a fixed point: mass exactly conserved
        assert abs(p.mean() - p1.mean()) < 1e-14 and abs(c.mean() - c1.mean()) < 1e-14
        e = s.energy()[0]
        assert e < e_prev
        e_prev = e


def _numpy_step(m, kl, km, c2a, c2b, phi, con):
    """One reference step (misfit.cpp:479-498) with numpy FFTs, narrowing off."""
    L, M = m.L, m.M
    P = L * M
    ks = (kl[:, None].astype(np.float32) ** 2 + km[None, :].astype(np.float32) ** 2).astype(np.float32)
    nu, xi, w, c0 = 1.399999976158142, 1.0, 0.019999999552965164, 0.5
    phat = np.fft.fft2(phi)
    a = np.fft.ifft2(c2a * phat).real
    b = np.fft.ifft2(c2b * phat).real
    h = lambda x: 1 - 3 * x ** 2 + 2 * x ** 3
    in3 = -h(con) * a - h(1 - con) * b
    inn = -nu * phi ** 2 / 2 + xi * phi ** 3 / 3 + w * (con * np.log(con / c0) + (1 - con) * np.log((1 - con) / (1 - c0))) + in3
    inc = (w * (phi + 1) * np.log(con / c0) - w * (phi + 1) * np.log((1 - con) / (1 - c0))
           - 3 * phi * (a - b) * con ** 2 - 3 * phi * (b - a) * con)
    outn, outc, chat = np.fft.fft2(inn), np.fft.fft2(inc), np.fft.fft2(con)
    mk = (np.float32(m.mn) * np.float32(m.dt) * ks).astype(np.float32).astype(np.float64)
    mkc = (np.float32(m.mc) * np.float32(m.dt) * ks).astype(np.float32).astype(np.float64)
    den = (np.float32(np.float32(m.mc) * np.float32(m.dt) * np.float32(m.alpha)) * ks * ks).astype(np.float32)
    phat = (phat - mk * outn) / (1 + mk)
    chat = (chat - mkc * outc) / (1 + den.astype(np.float64))
    return np.fft.ifft2(phat).real, np.fft.ifft2(chat).real


@pytest.mark.parametrize("kind", [0, 1])
def test_step_matches_numpy_restatement(oracle, kind):
    lib = oracle.load()
    L = M = 32
    kw = dict(grainx=2, grainy=2) if kind == 0 else {}
    m = oracle.model(kind, L, M, 1, narrow=0, **kw)
    kl, km, _ = _ktabs(oracle, m)
    c2a = np.zeros(L * M, np.float32)
    c2b = np.zeros(L * M, np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.pfo_c2(ctypes.byref(m), c2a.ctypes.data_as(fp), c2b.ctypes.data_as(fp), None, None)
    phi, con = oracle.init_fields(m)
    s = oracle.Sim(m, phi, con)
    p, c = phi.reshape(L, M), con.reshape(L, M)
    for _ in range(5):
        p, c = _numpy_step(m, kl, km, c2a.reshape(L, M).astype(np.float64), c2b.reshape(L, M).astype(np.float64), p, c)
    s.run(5)
    op, oc = s.fields()
    assert np.abs(op - p.ravel()).max() < 1e-12
    assert np.abs(oc - c.ravel()).max() < 1e-12


# --------------------------------------------------------------------------
# Phimax / strain on planted fields
# --------------------------------------------------------------------------
def _bumps(L, M, centres, width=0.3, dx=0.125):
    x = (np.arange(L) * dx)[:, None]
    y = (np.arange(M) * dx)[None, :]
    f = np.full((L, M), -0.5)
    for cx, cy, amp in centres:
        f += amp * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * width ** 2))
    return f


def test_phimax_misfit_quadratic_peaks(oracle):
    """An exact paraboloid around each maximum is recovered exactly by the
    quadratic sub-pixel fit: x = (i+2)dx + delta, y = (m+1)dx + delta."""
    L = M = 32
    dx = 0.125
    f = np.full((L, M), -1.0)
    peaks = [(10, 7, 0.031, -0.02, 0.9), (20, 25, -0.04, 0.05, 0.7)]
    for i0, m0, ox, oy, top in peaks:
        for di in range(-2, 3):
            for dm in range(-2, 3):
                X, Y = di * dx - ox, dm * dx - oy
                f[i0 + di, m0 + dm] = top - 4 * X * X - 3 * Y * Y
    x, y, z, ph, co = oracle.phimax_misfit(L, M, 1, dx, 0.2, f.ravel(), np.full(L * M, 0.3))
    assert len(x) == 2
    # scan order is l fastest within m: the peak with smaller m comes first
    for (i0, m0, ox, oy, top), xi, yi, vi in zip(sorted(peaks, key=lambda q: q[1]), x, y, ph):
        assert abs(xi - ((i0 + 2) * dx + ox)) < 1e-12
        assert abs(yi - ((m0 + 1) * dx + oy)) < 1e-12
        assert abs(vi - top) < 1e-12
    assert np.all(z == 0.125) and np.all(co == np.float32(0.3))


def test_phimax_misfit_fallback_and_edges(oracle):
    L = M = 16
    f = np.full((L, M), -0.5)             # nothing above the -0.07 cut
    f[3, 4] = -0.2
    x, y, z, ph, co = oracle.phimax_misfit(L, M, 1, 0.125, 0.2, f.ravel(), np.full(L * M, 0.25))
    assert len(x) == 1 and x[0] == (3 + 2) * 0.125 and y[0] == 5 * 0.125 and ph[0] == -0.2
    # a maximum on the first row: its x halo is the row itself (stub Sendrecv)
    f = np.full((L, M), 0.0)
    f[0, 8] = 1.0
    x, *_ = oracle.phimax_misfit(L, M, 1, 0.125, 0.2, f.ravel(), np.full(L * M, 0.25))
    assert len(x) == 1


def test_phimax_cvd_grid_maxima(oracle):
    L = M = 32
    f = _bumps(L, M, [(1.0, 1.5, 1.0), (2.5, 3.0, 0.8)]).ravel()
    px, py, pz, pp, phic, pmax = oracle.phimax_cvd(L, M, 1, 0.125, 0.01, f)
    assert pmax == np.float32(f.max())
    assert phic == np.float32(np.float32(0.01) + (np.float32(f.max()) - np.float32(0.01)) * np.float32(0.618))
    assert len(px) >= 1
    # first maximum: l = 8 (x = 1.0), m = 12 (y = 1.5); pc_z = (l+1) dx / Lx
    assert pz[0] == np.float32(np.float32(9 * 0.125) / np.float32(32 * 0.125))
    assert px[0] == np.float32(np.float32(13 * 0.125) / np.float32(32 * 0.125))


def test_phimax_elastic_peaks(oracle):
    L = M = 32
    f = _bumps(L, M, [(1.0, 1.5, 1.0), (2.5, 3.0, 0.9)]).astype(np.float32)
    x, y, z, ph = oracle.phimax_elastic(L, M, 1, 0.125, f.ravel())
    assert len(x) == 2
    # symmetric bumps centred on grid points: no sub-pixel shift beyond the
    # +1e-7 regularisers; z = dx - 0.5 for N = 1 (delta_z = -1e-7/2e-7)
    assert abs(x[0] - (8 + 2) * 0.125) < 1e-5 and abs(y[0] - 13 * 0.125) < 1e-5
    assert np.allclose(z, 0.125 - 0.5)


def test_strain_small_cases(oracle):
    L = M = 32
    dx = 0.125
    # square lattice a = 1: neighbours at 1 (4) and sqrt(2) (4), zero strain
    g = np.array([(i, j) for i in range(4) for j in range(4)], float)
    s, nb = oracle.strain(L, M, 1, dx, 1.7, g[:, 0], g[:, 1], np.zeros(16))
    assert np.all(nb[5:7] == 8) and np.allclose(s, 0.0)
    # chain: left neighbour at 1.0, right at 1.2 -> |strain| = 0.2 at atom 1
    x = np.array([0.0, 1.0, 2.2])
    s, nb = oracle.strain(L, M, 1, dx, 1.7, x, np.zeros(3), np.zeros(3))
    assert nb.tolist() == [1, 2, 1] and abs(s[1] - 0.2) < 1e-6 and abs(s[0]) < 1e-7
    # minimum image across the periodic box (Lx = 4)
    x = np.array([0.1, 3.3])
    s, nb = oracle.strain(32, 32, 1, dx, 1.7, x, np.zeros(2), np.zeros(2))
    assert nb.tolist() == [1, 1] and abs(s[0] - 0.2) < 1e-6
    # more than 16 neighbours: all counted, only the first 16 (j order) summed
    ang = np.linspace(0, 2 * np.pi, 20, endpoint=False)
    x = np.concatenate([[10.0], 10 + 1.1 * np.cos(ang)])
    y = np.concatenate([[10.0], 10 + 1.1 * np.sin(ang)])
    s, nb = oracle.strain(L, M, 1, dx, 1.7, x, y, np.zeros(21))
    assert nb[0] == 20
    vx = 0.1 * np.cos(ang[:16]).sum()
    vy = 0.1 * np.sin(ang[:16]).sum()
    assert abs(s[0] - np.hypot(vx, vy)) < 1e-6


# --------------------------------------------------------------------------
# VTI format
# --------------------------------------------------------------------------
def test_vti_roundtrip_and_layout(oracle, tmp_path):
    lib = oracle.load()
    L, M, N = 5, 3, 2
    f = np.arange(L * M * N, dtype=np.float64) * 0.5
    p = str(tmp_path / "phi_7_0.vti").encode()
    lib.pfo_write_vti(p, b"phi", L, M, N, f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    raw = open(p, "rb").read()
    head, rest = raw.split(b"_", 1) if raw.count(b"\n_") == 0 else raw.split(b"\n_", 1)
    assert head.count(b"\n") in (11, 12)
    nbytes = int(np.frombuffer(rest[:4], np.int32)[0])
    assert nbytes == L * M * N * 4 + 4
    payload = np.frombuffer(rest[4:4 + 4 * L * M * N], np.float32)
    # payload order: n outer, m, l inner (l fastest) of index n + N*(m + M*l)
    ref = f.reshape(L, M, N).transpose(2, 1, 0).ravel().astype(np.float32)
    assert np.array_equal(payload, ref)
    assert rest[4 + 4 * L * M * N:] == b"\n</AppendedData>\n</VTKFile>\n"
    back = np.zeros(L * M * N, np.float32)
    n = lib.pfo_read_vti(p, L, M, N, back.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert n == L * M * N and np.array_equal(back, f.astype(np.float32))


# ---------------------------------------------------------------------------
# cryst_order (elastic.cpp:3013-3389): Lechner-Dellago averaged q_l.
# Pinned against textbook Steinhardt values computed here with scipy's
# spherical harmonics on perfect lattices (every atom equivalent, so the
# neighbour average equals the atom's own q_lm).  The reference's "q4"
# column is l = 6 (:3155) and its q6 column keeps adding onto the same
# per-atom sum, i.e. sqrt(2) times it.
# ---------------------------------------------------------------------------
def _textbook_q(l, vecs):
    """Steinhardt q_l of one neighbour shell, with the reference's angles:
    polar = acos(dz/(r + 1e-4)), azimuth = atan(dy/(dx + 1e-4)) (+pi for
    dx < 0) (elastic.cpp:3099-3104)."""
    import scipy.special as sps
    r = np.linalg.norm(vecs, axis=1)
    polar = np.arccos(vecs[:, 2] / (r + 1e-4))
    azim = np.arctan(vecs[:, 1] / (vecs[:, 0] + 1e-4)) + np.where(vecs[:, 0] < 0, 3.1415926, 0.0)
    tot = 0.0
    for m in range(-l, l + 1):
        if hasattr(sps, "sph_harm_y"):
            y = sps.sph_harm_y(l, m, polar, azim)
        else:
            y = sps.sph_harm(m, l, azim, polar)
        tot += abs(y.mean()) ** 2
    return np.sqrt(4 * np.pi / (2 * l + 1) * tot)


def _lattice(kind, cells, a, off=0.137):
    basis = {"sc": [(0, 0, 0)], "bcc": [(0, 0, 0), (0.5, 0.5, 0.5)],
             "fcc": [(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)]}[kind]
    pts = [((i + bx) * a + off, (j + by) * a + off, (k + bz) * a + off)
           for i in range(cells) for j in range(cells) for k in range(cells) for bx, by, bz in basis]
    return np.array(pts)


@pytest.mark.parametrize("kind,a,nn", [("sc", 1.0, 6), ("bcc", 1.3, 8), ("fcc", 2 ** 0.5, 12)])
def test_cryst_order_lattices(oracle, kind, a, nn):
    cells, L = 4, 16
    dx = float(np.float32(cells * a / L))  # box = L*dx = cells*a
    pts = _lattice(kind, cells, a)
    q4, q6, nb = oracle.cryst_order(L, L, L, dx, pts[:, 0], pts[:, 1], pts[:, 2])
    first = {"sc": [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)],
             "bcc": [(sx * .5, sy * .5, sz * .5) for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)],
             "fcc": [v for v in [(sx * .5, sy * .5, 0) for sx in (1, -1) for sy in (1, -1)] +
                     [(sx * .5, 0, sz * .5) for sx in (1, -1) for sz in (1, -1)] +
                     [(0, sy * .5, sz * .5) for sy in (1, -1) for sz in (1, -1)]]}[kind]
    want = _textbook_q(6, np.array(first) * a)
    assert np.all(nb == nn)
    assert np.abs(q4 - want).max() < 2e-5, (q4[:3], want)
    assert np.abs(q6 - np.sqrt(2.0) * q4).max() < 1e-6


def test_cryst_order_cap_and_isolated(oracle):
    # a dense blob: every atom has > 16 neighbours within 1.2 -> q = 0;
    # one far atom without neighbours -> q = 0 and Nb = 0.
    g = np.linspace(0.0, 0.9, 3)
    blob = np.array([(x + 2, y + 2, z + 2) for x in g for y in g for z in g])
    pts = np.vstack([blob, [[6.0, 6.0, 6.0]]])
    q4, q6, nb = oracle.cryst_order(32, 32, 32, 0.25, pts[:, 0], pts[:, 1], pts[:, 2])
    assert nb[-1] == 0 and q4[-1] == 0 and q6[-1] == 0
    assert nb[13] == 26 and q4[13] == 0 and q6[13] == 0  # the blob's centre


def test_ylm_reference_conventions(oracle):
    """The reference's Ylm: the Condon-Shortley phase twice (in P_l^m and in
    the (-1)^m normalisation), so (-1)^m times the textbook Y_lm for m >= 0;
    negative m reuses |m| in the real part and drops the (-1)^m sign in the
    imaginary part.  |Y|^2 and hence q_l are unaffected."""
    import scipy.special as sps
    th, fa = 0.7, 1.1
    for m in range(0, 7):
        yr, yc = oracle.ylm(6, m, th, fa)
        if hasattr(sps, "sph_harm_y"):
            t = sps.sph_harm_y(6, m, th, fa)
        else:
            t = sps.sph_harm(m, 6, fa, th)
        sgn = (-1) ** m
        assert abs(yr - sgn * t.real) < 2e-6 and abs(yc - sgn * t.imag) < 2e-6, (m, yr, yc, t)
        yr2, yc2 = oracle.ylm(6, -m, th, fa)
        assert yr2 == yr
        assert yc2 == pytest.approx(-(-1) ** m * yc, abs=1e-7)


# ---------------------------------------------------------------------------
# CVD GR / Q6_GLO (cvd.cpp:1277-1510)
# ---------------------------------------------------------------------------
def _frac_lattice(kind, cells, a, box, off=0.01):
    return (_lattice(kind, cells, a, off=off * a) / box).astype(np.float32)


def test_cvd_gr_sc_shells(oracle):
    """Simple cubic, spacing 1: the histogram holds n times the coordination
    number of each shell (6, 12, 8, 6, 24, 24 at 1, sqrt2, sqrt3, 2, sqrt5,
    sqrt6), 2 per pair; nothing inside a/sqrt(2)/2."""
    L, dx = 64, 0.125                       # box 8 = 8 cells
    f = _frac_lattice("sc", 8, 1.0, 8.0)
    h = oracle.cvd_gr(L, L, L, dx, f[:, 0], f[:, 1], f[:, 2])
    n = len(f)
    assert len(h) == int(np.float32(L * dx) / np.float32(0.04)) == 200
    want = {int(np.sqrt(d2) / 0.04): c for d2, c in ((1, 6), (2, 12), (3, 8), (4, 6), (5, 24), (6, 24))}
    for b, c in want.items():
        assert h[b] == c * n, (b, h[b], c * n)
    assert h[:25].sum() == 0


def test_cvd_gr_random_bruteforce(oracle):
    """Random fractions against a float64 minimum-image pair count: bins may
    differ only by pairs within 1e-4 of a bin edge."""
    rng = np.random.default_rng(5)
    L, M, N, dx = 48, 48, 16, 0.125
    f = rng.random((300, 3)).astype(np.float32)
    h = oracle.cvd_gr(L, M, N, dx, f[:, 0], f[:, 1], f[:, 2]).astype(np.int64)
    fr = np.array([L, M, N], np.float64) * dx
    d = (f[:, None, :].astype(np.float64) - f[None, :, :]) * fr
    d = np.minimum(np.abs(d), np.minimum(np.abs(d + fr), np.abs(d - fr)))
    r = np.sqrt((d ** 2).sum(-1)) + 1e-6
    iu = np.triu_indices(len(f), 1)
    r = r[iu]
    b = np.floor(r / 0.04).astype(np.int64)
    keep = (r > 1 / np.sqrt(2) / 2) & (b < len(h))
    want = 2 * np.bincount(b[keep], minlength=len(h))[:len(h)]
    edge = np.abs(r / 0.04 - np.round(r / 0.04)) < 1e-4
    assert np.abs(h - want).sum() <= 4 * edge.sum()
    assert h.sum() > 0


def test_cvd_gr_files(oracle, tmp_path):
    """Grq6: nideal telescopes to 4/3 pi rho R^3 and gr * nideal * num_t
    gives back the histogram."""
    rng = np.random.default_rng(6)
    L, dx = 64, 0.125
    f = rng.random((200, 3)).astype(np.float32)
    h = oracle.cvd_gr(L, L, L, dx, f[:, 0], f[:, 1], f[:, 2])
    lib = oracle.load()
    import ctypes
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        lib.pfo_write_cvd_gr(40, L, L, L, ctypes.c_float(dx), len(f),
                             f[:, 0].copy().ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                             h.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    finally:
        os.chdir(cwd)
    hist = np.loadtxt(tmp_path / "hist_40.txt")
    assert hist.shape == (200, 3) and np.array_equal(hist[:, 1], h) and (hist[:, 2] == 200).all()
    g = np.loadtxt(tmp_path / "Grq6_40.txt")
    rho = 200 / 8.0 ** 3
    assert abs(g[:, 3].sum() - 4 / 3 * np.pi * rho * (200 * 0.04) ** 3) < 1e-2
    # gr * nideal * num_t = hist, up to the 4-decimal rounding of gr and nideal
    err = np.abs(g[:, 1] * g[:, 3] * 200 - g[:, 2])
    assert (err <= 200 * 5e-5 * (g[:, 1] + g[:, 3] + 1e-4) + 1e-9).all()
    assert np.allclose(g[:, 0], np.arange(200) * 0.04, atol=6e-5)


def _textbook_q6_glo(vecs_list):
    """Local Steinhardt q6 per atom and the bond-weighted global Q6 with the
    angles of Q6_GLO: polar acos(dz/(r + 1e-6)), azimuth atan(dy/dx) without
    the quadrant term, 0 when dy == 0 (cvd.cpp:1414-1422)."""
    import scipy.special as sps

    def y6(v):
        r = np.linalg.norm(v, axis=1) + 1e-6
        polar = np.arccos(v[:, 2] / r)
        with np.errstate(divide="ignore"):
            azim = np.where(v[:, 1] == 0, 0.0, np.arctan(v[:, 1] / v[:, 0]))
        out = []
        for m in range(-6, 7):
            out.append(sps.sph_harm_y(6, m, polar, azim) if hasattr(sps, "sph_harm_y")
                       else sps.sph_harm(m, 6, azim, polar))
        return np.array(out)            # [13, nb]

    loc, tot, nbond = [], np.zeros(13, complex), 0
    for v in vecs_list:
        y = y6(v)
        loc.append(np.sqrt(4 * np.pi / 13 * (np.abs(y.mean(1)) ** 2).sum()))
        tot += y.sum(1)
        nbond += len(v)
    return np.array(loc), np.sqrt(4 * np.pi / 13 * (np.abs(tot / nbond) ** 2).sum())


@pytest.mark.parametrize("kind,a,nn", [("sc", 0.9, 6), ("bcc", 1.0, 14), ("fcc", 0.9 * 2 ** 0.5, 12)])
def test_cvd_q6_lattices(oracle, kind, a, nn):
    """Perfect lattices in a periodic box: the window a*sqrt(3)/2*[0.9,1.2]
    takes 6 (sc), 8+6 (bcc, a = 1) and 12 (fcc) neighbours; local and global
    Q6 equal the textbook values (0.3536, 0.5107, 0.5745)."""
    cells = 4
    L = 32
    box = cells * a
    dx = float(np.float32(box / L))
    f = _frac_lattice(kind, cells, a, L * np.float32(dx))
    nb, ql, qg, hq = oracle.cvd_q6(L, L, L, dx, f[:, 0], f[:, 1], f[:, 2])
    want = {"sc": 0.35355, "bcc": 0.51069, "fcc": 0.57452}[kind]
    assert (nb == nn).all()
    assert np.abs(ql - want).max() < 5e-5 and abs(qg - want) < 5e-5
    assert hq[int(want / 0.05)] == len(f) and hq.sum() == len(f)


def test_cvd_q6_jitter_textbook(oracle):
    """Jittered fcc: per-atom local Q6 and the global Q6 against scipy's
    Y_6m over the neighbours found by a float64 brute-force search."""
    rng = np.random.default_rng(11)
    a, cells, L = 0.9 * 2 ** 0.5, 4, 32
    box = cells * a
    dx = float(np.float32(box / L))
    pts = _lattice("fcc", cells, a, off=0.05) + 0.03 * rng.standard_normal((4 * cells ** 3, 3))
    fb = np.float32(L * np.float32(dx))
    f = (pts / fb).astype(np.float32)
    nb, ql, qg, hq = oracle.cvd_q6(L, L, L, dx, f[:, 0], f[:, 1], f[:, 2])
    p = f.astype(np.float64) * float(fb)
    vecs = []
    for i in range(len(p)):
        d = p - p[i]
        d = d - float(fb) * np.round(d / float(fb))
        r = np.sqrt((d ** 2).sum(1)) + 1e-6
        sel = (r <= np.sqrt(3) / 2 * 1.2 + 1e-3) & (r >= np.sqrt(3) / 2 * 0.9 - 1e-3)
        sel[i] = False
        vecs.append(d[sel])
    assert np.array_equal(nb, [len(v) for v in vecs])
    loc, glob = _textbook_q6_glo(vecs)
    assert np.abs(ql - loc).max() < 2e-5
    assert abs(qg - glob) < 2e-5
    assert 0.3 < ql.mean() < 0.58


def test_cvd_q6_cap(oracle):
    """More than 24 neighbours: local Q6 = 0, yet the count enters the global
    normalisation sum_i Nb_i."""
    import scipy.special as sps
    rng = np.random.default_rng(2)
    L, dx = 32, 0.25
    fb = L * dx
    p = rng.random((400, 3)) * 3.0 + 2.0
    f = (p / fb).astype(np.float32)
    nb, ql, qg, hq = oracle.cvd_q6(L, L, L, dx, f[:, 0], f[:, 1], f[:, 2])
    assert nb.max() > 24 and nb.min() <= 24
    assert (ql[nb > 24] == 0).all() and (ql[(nb > 0) & (nb <= 24)] > 0).all()
    p = f.astype(np.float64) * fb
    tot = np.zeros(13, complex)
    for i in range(len(p)):
        if not 0 < nb[i] <= 24:
            continue
        d = p - p[i]
        r = np.sqrt((d ** 2).sum(1)) + 1e-6
        sel = (r <= np.sqrt(3) / 2 * 1.2 + 1e-3) & (r >= np.sqrt(3) / 2 * 0.9 - 1e-3)
        sel[i] = False
        v = d[sel]
        polar = np.arccos(v[:, 2] / r[sel])
        azim = np.where(v[:, 1] == 0, 0.0, np.arctan(v[:, 1] / v[:, 0]))
        for k, m in enumerate(range(-6, 7)):
            y = (sps.sph_harm_y(6, m, polar, azim) if hasattr(sps, "sph_harm_y")
                 else sps.sph_harm(m, 6, azim, polar))
            tot[k] += y.sum()
    want = np.sqrt(4 * np.pi / 13 * (np.abs(tot / nb.astype(np.float64).sum()) ** 2).sum())
    assert abs(qg - want) < 2e-5, (qg, want)
