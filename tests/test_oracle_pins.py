"""Pins of the oracle's ENERGYC2 and of the Misfit grain lattice by what the
mathematics fixes, independently of the oracle's own formulas.

ENERGYC2 (misfit.cpp:1043-1092; cvd.cpp:979-1040) on states whose energy has
a closed form:
  * a uniform state: phi_hat and con_hat are pure k = 0, so a = C2A(0) phi0,
    b = C2B(0) phi0 and the gradient vanishes; the energy density is the same
    at every point and is restated here from the reference's expressions
    (constants of misfit.cpp:122-136, cvd.cpp:39-45);
  * a single cosine in phi at a wave number where C2A and C2B are large
    (C2 of a uniform state, e^{-(2 pi)^2/1.28}, is ~1e-14, so only a
    modulated phi exercises the coupling term phi/2 * in3): a and b are the
    kernel values times the cosine, restated from misfit.cpp:714, :743, 826;
  * a single cosine in con along x (l), y (m) and, on a 3-D grid, z (n): the
    gradient of A cos(k x) is -A k sin(k x) exactly, and UPDATEINC2GRA sums
    g1^2 + g2^2 + g2^2 (misfit.cpp:915: the y component twice, z never), so
    the y cosine carries twice the x cosine's gradient energy and the z
    cosine none.  A dropped or doubled term, a wrong coefficient, or the
    textbook g1^2 + g2^2 + g3^2 fails one of these.
The grain lattice of INITIAL_dis (misfit.cpp:645-676): a patch is a square
lattice of wave number (1 - con0) ka0 + con0 kb0 rotated by the grain's
angle theta, mirrored (-theta) on grains with odd i + j + k (the cs sign of
misfit.cpp:660); its structure factor peaks at that angle (modulo 90 degrees,
the lattice's symmetry).  A sign slip in the rotated X or Y moves the peak to
90 - theta.
"""
import ctypes

import numpy as np
import pytest

F = np.float32
NIU, KSI, C0, ALPHA, W = F(1.4), F(1.0), F(0.5), F(0.5), F(0.02)


def _c2_at_zero(m, which):
    """C2A(0) / C2B(0): the |k| <= kc branch of misfit.cpp:714 / :743 at k = 0."""
    if which == "a":
        sg, g, k1 = F(m.sigMa1), F(m.gaa1), F(m.ka1)
    else:
        sg, g, k1 = F(m.sigMb1), F(m.gab1), F(m.kb1)
    s = F(F(m.sig) / sg)
    return F(np.exp(-float(s) ** 2) * np.exp(-float(F(0.0) - k1) ** 2 / float(F(2) * g * g)))


def _fmix(c):
    c0 = float(C0)
    return c * np.log(c / c0) + (1.0 - c) * np.log((1.0 - c) / (1.0 - c0))


def _h(c):
    return 1.0 - 3.0 * c ** 2 + 2.0 * c ** 3


def _density(phi, c, a, b, grac, narrow):
    """ENERGYC2's per-point sum energy_id + energy_mix + energy_ex; with the
    reference's float temporaries when narrow."""
    eid = phi ** 2 / 2.0 - float(NIU) * phi ** 3 / 6.0 + float(KSI) * phi ** 4 / 12.0
    emix = (phi + 1.0) * float(W) * _fmix(c)
    in3 = -_h(c) * a - _h(1.0 - c) * b
    eex = 0.5 * phi * in3 + float(ALPHA) * grac
    if narrow:
        return (F(eid) + F(emix)) + F(eex)
    return eid + emix + eex


def _sim_energy(oracle, m, phi, con):
    s = oracle.Sim(m, phi, con)
    return s.energy()


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("narrow", [0, 1])
def test_energy_uniform_state(oracle, kind, narrow):
    m = oracle.model(kind, 64, 32, 1, narrow=narrow)
    assert (F(m.niu), F(m.ksi), F(m.c0), F(m.alpha), F(m.w)) == (NIU, KSI, C0, ALPHA, W)
    P = 64 * 32
    for phi0, c in ((0.23, 0.31), (-0.4, 0.72), (0.05, 0.2)):
        a = float(_c2_at_zero(m, "a")) * phi0
        b = float(_c2_at_zero(m, "b")) * phi0
        want = float(_density(phi0, c, a, b, 0.0, narrow))
        e, rho = _sim_energy(oracle, m, np.full(P, phi0), np.full(P, c))
        assert abs(e - want) <= 1e-12 * abs(want) + 1e-16, (phi0, c, e, want)
        assert abs(rho - phi0) <= 1e-12  # P sequential additions of phi0


def _c2_at(m, which, k):
    """C2A / C2B at |k| (misfit.cpp:700-750): float kc root, float sqrt(ks)."""
    if which == "a":
        sig1, sig2, g1, g2, k1, k2 = (F(v) for v in (m.sigMa1, m.sigMa2, m.gaa1, m.gaa2, m.ka1, m.ka2))
    else:
        sig1, sig2, g1, g2, k1, k2 = (F(v) for v in (m.sigMb1, m.sigMb2, m.gab1, m.gab2, m.kb1, m.kb2))
    sig = F(m.sig)
    aa1, aa2 = F(float(F(sig / sig1)) ** 2), F(float(F(sig / sig2)) ** 2)
    bb1, bb2 = F(1.0 / (2.0 * float(g1) * float(g1))), F(1.0 / (2.0 * float(g2) * float(g2)))
    t1 = F(bb1 * k1 - bb2 * k2)
    t2 = F((bb1 - bb2) * F(F(aa1 - aa2) + F(bb1 * k1 * k1) - F(bb2 * k2 * k2)))
    kc = F((float(t1) + np.sqrt(float(t1) ** 2 - float(t2))) / float(bb1 - bb2))
    sk = F(np.sqrt(F(F(k) * F(k))))
    s1, kk, g = (sig1, k1, g1) if sk <= kc else (sig2, k2, g2)
    return F(np.exp(-float(F(sig / s1)) ** 2) * np.exp(-float(F(sk - kk)) ** 2 / float(F(2) * g * g)))


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("narrow", [0, 1])
def test_energy_coupling_term(oracle, kind, narrow):
    """phi = phi0 + B cos(k x) with k = k_l[2] = 2 pi (float pi) on 16 x 8,
    near the C2A peak; con uniform.  a(x) = C2A(0) phi0 + C2A(k) B cos(k x),
    b likewise, in3 = -h(c) a - h(1-c) b (misfit.cpp:826), energy_ex =
    phi in3 / 2 (misfit.cpp:1056)."""
    L, M = 16, 8
    m = oracle.model(kind, L, M, 1, narrow=narrow)
    kl, _, _ = _ktable(oracle, m)
    k = float(kl[2])
    ca, cb = float(_c2_at(m, "a", k)), float(_c2_at(m, "b", k))
    assert ca > 0.9 and ca - cb > 0.1  # the coupling term is large and a, b differ
    ca0, cb0 = float(_c2_at_zero(m, "a")), float(_c2_at_zero(m, "b"))
    x = np.arange(L)[:, None] * np.ones((1, M))
    for phi0, B, c in ((0.1, 0.3, 0.27), (-0.2, 0.45, 0.64)):
        cosx = np.cos(2.0 * np.pi * 2 * x / L)
        phi = phi0 + B * cosx
        a = ca0 * phi0 + ca * B * cosx
        b = cb0 * phi0 + cb * B * cosx
        want = float(np.mean(_density(phi, c, a, b, 0.0, narrow)))
        e, _ = _sim_energy(oracle, m, phi.ravel(), np.full(L * M, c))
        tol = 1e-6 if narrow else 1e-12  # float rounding of near-equal doubles may differ by an ulp
        assert abs(e - want) <= tol * abs(want), (phi0, B, c, e, want)


def _ktable(oracle, m):
    lib = oracle.load()
    kl = np.zeros(m.L, np.float32)
    km = np.zeros(m.M, np.float32)
    kn = np.zeros(m.N, np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.pfo_ktables(ctypes.byref(m), kl.ctypes.data_as(fp), km.ctypes.data_as(fp), kn.ctypes.data_as(fp))
    return kl, km, kn


@pytest.mark.parametrize("kind", [0, 1])
def test_energy_gradient_term_weights(oracle, kind):
    """con = c1 + A cos(2 pi q i / n) along one axis at a time, phi uniform:
    x carries alpha (A k sin)^2, y twice that, z nothing (misfit.cpp:915)."""
    L = M = N = 16
    q, c1, A, phi0 = 3, 0.35, 0.12, 0.17
    m = oracle.model(kind, L, M, N, narrow=0)
    kl, km, kn = _ktable(oracle, m)
    assert kl[q] == km[q] == kn[q]
    k = float(kl[q])
    a = float(_c2_at_zero(m, "a")) * phi0
    b = float(_c2_at_zero(m, "b")) * phi0
    idx = np.indices((L, M, N))
    got, want = {}, {}
    for axis, weight in ((0, 1.0), (1, 2.0), (2, 0.0)):
        th = 2.0 * np.pi * q * idx[axis] / (L, M, N)[axis]
        c = c1 + A * np.cos(th)
        grac = weight * (A * k * np.sin(th)) ** 2
        want[axis] = float(np.mean(_density(phi0, c, a, b, grac, 0)))
        got[axis], _ = _sim_energy(oracle, m, np.full(L * M * N, phi0), c.ravel())
        assert abs(got[axis] - want[axis]) <= 1e-12 * abs(want[axis]), (axis, got[axis], want[axis])
    # the y cosine carries alpha A^2 k^2 / 2 more than the x cosine, which
    # carries as much more than the z cosine (same local terms in all three)
    gx = float(ALPHA) * (A * k) ** 2 / 2.0
    assert abs((got[1] - got[0]) - gx) <= 1e-9 * gx
    assert abs((got[0] - got[2]) - gx) <= 1e-9 * gx


def _structure_peak(f, kk, dx):
    L, M = f.shape
    l = np.arange(L)[:, None]
    mm = np.arange(M)[None, :]
    angs = np.radians(np.arange(0.0, 90.0, 0.05))
    s = np.array([abs(np.sum(f * np.exp(-1j * kk * dx * (l * np.cos(t) + mm * np.sin(t))))) for t in angs])
    return float(np.degrees(angs[s.argmax()])), s


def test_misfit_grain_lattice_orientation(oracle):
    lib = oracle.load()
    L = M = 128
    m = oracle.model(0, L, M, 1, grainx=2, grainy=1)
    phi, _ = oracle.init_fields(m)
    phi = phi.reshape(L, M)
    dx = float(F(m.dx))
    kk = float((F(1) - F(m.con0)) * F(m.ka0) + F(m.con0) * F(m.kb0))
    l = np.arange(L)[:, None]
    mm = np.arange(M)[None, :]
    for g in range(2):
        theta = np.degrees(lib.pfo_grain_theta(ctypes.byref(m), g, 0, 0))
        cx, cy = m.dnx * g + m.dnx // 2, m.dny // 2
        patch = (np.abs((l - cx) * dx) < L * dx / 2 / 2 - 0.2) & (np.abs((mm - cy) * dx) < M * dx / 2 - 0.2)
        f = np.where(patch, phi - float(F(m.u0)), 0.0)
        peak, s = _structure_peak(f, kk, dx)
        expect = (theta if g % 2 == 0 else -theta) % 90.0
        dist = min(abs(peak - expect), 90.0 - abs(peak - expect))
        assert dist < 0.15, (g, theta, peak)
        assert s.max() > 20 * np.median(s)
        # the mirrored angle is far off the peak
        mirror = (-expect) % 90.0
        assert min(abs(peak - mirror), 90.0 - abs(peak - mirror)) > 5.0
    # outside the patches (|y - cy| dx >= M dx / 2 - 0.2) the field is u0
    assert np.all(phi[:, M - 1] == F(m.u0))
