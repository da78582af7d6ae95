"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the parity checker.  Never used by the
product path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
CLI_PATH = os.path.join(HERE, "pfc-oracle-cli")

MISFIT, CVD = 0, 1


class Model(ctypes.Structure):
    _fields_ = (
        [("model", ctypes.c_int), ("L", ctypes.c_int), ("M", ctypes.c_int), ("N", ctypes.c_int)]
        + [(n, ctypes.c_float) for n in "u0 sig con0 dt dx axx b0".split()]
        + [(n, ctypes.c_int) for n in "grainx grainy grainz dnx dny dnz".split()]
        + [("seed", ctypes.c_uint), ("steps", ctypes.c_int), ("mod", ctypes.c_int)]
        + [(n, ctypes.c_float) for n in ("a_lattice b_lattice gaa1 gaa2 gab1 gab2 niu ksi ka0 ka1 ka2 sigMa1 "
                                         "sigMa2 kb0 kb1 kb2 sigMb1 sigMb2 c0 mn mc alpha w").split()]
        + [("narrow", ctypes.c_int)]
    )


class Atoms(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("cap", ctypes.c_int),
                ("x", ctypes.POINTER(ctypes.c_double)), ("y", ctypes.POINTER(ctypes.c_double)),
                ("z", ctypes.POINTER(ctypes.c_double)), ("phi", ctypes.POINTER(ctypes.c_double)),
                ("con", ctypes.POINTER(ctypes.c_float))]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        lib = ctypes.CDLL(LIB_PATH)
        lib.pfo_splitmix64.restype = ctypes.c_uint64
        lib.pfo_rng_u01.restype = ctypes.c_double
        lib.pfo_gauss.restype = ctypes.c_uint64
        lib.pfo_gauss.argtypes = [ctypes.c_uint, ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)]
        lib.pfo_grain_theta.restype = ctypes.c_float
        lib.pfo_create.restype = ctypes.c_void_p
        lib.pfo_create.argtypes = [ctypes.POINTER(Model)]
        for f in ("pfo_destroy", "pfo_set_fields", "pfo_get_fields", "pfo_step", "pfo_energy", "pfo_run"):
            getattr(lib, f).restype = None
        lib.pfo_destroy.argtypes = [ctypes.c_void_p]
        dp = ctypes.POINTER(ctypes.c_double)
        lib.pfo_set_fields.argtypes = [ctypes.c_void_p, dp, dp]
        lib.pfo_get_fields.argtypes = [ctypes.c_void_p, dp, dp]
        lib.pfo_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.pfo_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.pfo_energy.argtypes = [ctypes.c_void_p, dp, dp]
        for f in ("pfo_in1", "pfo_in2", "pfo_in3", "pfo_inc"):
            getattr(lib, f).restype = ctypes.c_double
        lib.pfo_in1.argtypes = [ctypes.POINTER(Model), ctypes.c_double]
        lib.pfo_in2.argtypes = [ctypes.POINTER(Model), ctypes.c_double]
        lib.pfo_in3.argtypes = [ctypes.c_double] * 3
        lib.pfo_inc.argtypes = [ctypes.POINTER(Model)] + [ctypes.c_double] * 4
        lib.pfo_update_phi.argtypes = [ctypes.POINTER(Model), ctypes.c_float, dp, dp]
        lib.pfo_update_con.argtypes = [ctypes.POINTER(Model), ctypes.c_float, dp, dp]
        lib.pfo_update_phi.restype = None
        lib.pfo_update_con.restype = None
        lib.pfo_phimax_misfit.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_float, dp, dp,
                                                                ctypes.POINTER(Atoms)]
        fp = ctypes.POINTER(ctypes.c_float)
        lib.pfo_phimax_cvd.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_float, dp, fp, fp, fp, fp,
                                                             ctypes.c_int, fp, fp]
        lib.pfo_phimax_elastic.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float, fp, ctypes.POINTER(Atoms)]
        lib.pfo_strain.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_float, ctypes.c_int, dp, dp, dp, dp,
                                                         ctypes.POINTER(ctypes.c_uint)]
        lib.pfo_cryst_order.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_int, dp, dp, dp, dp, dp,
                                                            ctypes.POINTER(ctypes.c_uint)]
        lib.pfo_cryst_order.restype = None
        fpp = ctypes.POINTER(ctypes.c_float)
        lib.pfo_cvd_gr_bins.argtypes = [ctypes.c_int, ctypes.c_float]
        lib.pfo_cvd_gr.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_int, fpp, fpp, fpp,
                                                       ctypes.POINTER(ctypes.c_uint)]
        lib.pfo_cvd_gr.restype = None
        lib.pfo_cvd_q6.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_int, fpp, fpp, fpp,
                                                       ctypes.POINTER(ctypes.c_uint), fpp, fpp,
                                                       ctypes.POINTER(ctypes.c_ulonglong)]
        lib.pfo_cvd_q6.restype = None
        for fn in (lib.pfo_ylm_r, lib.pfo_ylm_c):
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float]
            fn.restype = ctypes.c_float
        lib.pfo_fft3d.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.pfo_fft1d.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        lib.pfo_ktables.argtypes = [ctypes.POINTER(Model), fp, fp, fp]
        lib.pfo_c2.argtypes = [ctypes.POINTER(Model), fp, fp, fp, fp]
        lib.pfo_init_fields.argtypes = [ctypes.POINTER(Model), dp, dp]
        lib.pfo_read_vti.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
        lib.pfo_write_vti.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def model(kind=MISFIT, L=64, M=64, N=1, **kw):
    lib = load()
    m = Model()
    lib.pfo_defaults(ctypes.byref(m), kind, L, M, N)
    for k, v in kw.items():
        setattr(m, k, v)
    lib.pfo_derive(ctypes.byref(m))
    return m


def init_fields(m):
    lib = load()
    P = m.L * m.M * m.N
    phi = np.zeros(P)
    con = np.zeros(P)
    lib.pfo_init_fields(ctypes.byref(m), _dp(phi), _dp(con))
    return phi, con


class Sim:
    def __init__(self, m, phi=None, con=None):
        self.lib = load()
        self.m = m
        self.P = m.L * m.M * m.N
        self.h = ctypes.c_void_p(self.lib.pfo_create(ctypes.byref(m)))
        if phi is None:
            phi, con = init_fields(m)
        self.set_fields(phi, con)
        self.t = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.pfo_destroy(self.h)
            self.h = None

    def set_fields(self, phi, con):
        phi = np.ascontiguousarray(phi, np.float64)
        con = np.ascontiguousarray(con, np.float64)
        self.lib.pfo_set_fields(self.h, _dp(phi), _dp(con))

    def run(self, n, faithful=0):
        self.lib.pfo_run(self.h, self.t, n, faithful)
        self.t += n

    def fields(self):
        phi = np.zeros(self.P)
        con = np.zeros(self.P)
        self.lib.pfo_get_fields(self.h, _dp(phi), _dp(con))
        return phi, con

    def energy(self):
        e = ctypes.c_double()
        r = ctypes.c_double()
        self.lib.pfo_energy(self.h, ctypes.byref(e), ctypes.byref(r))
        return e.value, r.value


def _atoms(cap):
    arrs = [np.zeros(cap) for _ in range(4)] + [np.zeros(cap, np.float32)]
    a = Atoms(0, cap, *[_dp(x) for x in arrs[:4]], _fp(arrs[4]))
    return a, arrs


def phimax_misfit(L, M, N, dx, con0, phi, con):
    lib = load()
    a, arrs = _atoms(L * M * max(N, 1))
    lib.pfo_phimax_misfit(L, M, N, ctypes.c_float(dx), ctypes.c_float(con0), _dp(np.ascontiguousarray(phi)),
                          _dp(np.ascontiguousarray(con)), ctypes.byref(a))
    return tuple(x[:a.n] for x in arrs)


def phimax_cvd(L, M, N, dx, u0, phi):
    lib = load()
    cap = L * M * N
    arrs = [np.zeros(cap, np.float32) for _ in range(4)]
    phic = ctypes.c_float()
    pmax = ctypes.c_float()
    n = lib.pfo_phimax_cvd(L, M, N, ctypes.c_float(dx), ctypes.c_float(u0), _dp(np.ascontiguousarray(phi)),
                           *[_fp(x) for x in arrs], cap, ctypes.byref(phic), ctypes.byref(pmax))
    return tuple(x[:n] for x in arrs) + (phic.value, pmax.value)


def phimax_elastic(L, M, N, dx, phi):
    lib = load()
    a, arrs = _atoms(L * M * N)
    lib.pfo_phimax_elastic(L, M, N, ctypes.c_float(dx), _fp(np.ascontiguousarray(phi, np.float32)), ctypes.byref(a))
    return tuple(x[:a.n] for x in arrs[:4])


def strain(L, M, N, dx, benchel, x, y, z):
    lib = load()
    n = len(x)
    s = np.zeros(n)
    nb = np.zeros(n, np.uint32)
    lib.pfo_strain(L, M, N, ctypes.c_float(dx), ctypes.c_float(benchel), n, _dp(np.ascontiguousarray(x)),
                   _dp(np.ascontiguousarray(y)), _dp(np.ascontiguousarray(z)), _dp(s),
                   nb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    return s, nb


def cryst_order(L, M, N, dx, x, y, z):
    """Per-atom (q4, q6, neighbour count) of cryst_order."""
    lib = load()
    n = len(x)
    q4 = np.zeros(n)
    q6 = np.zeros(n)
    nb = np.zeros(n, np.uint32)
    lib.pfo_cryst_order(L, M, N, ctypes.c_float(dx), n, _dp(np.ascontiguousarray(x, np.float64)),
                        _dp(np.ascontiguousarray(y, np.float64)), _dp(np.ascontiguousarray(z, np.float64)),
                        _dp(q4), _dp(q6), nb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    return q4, q6, nb


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def cvd_gr(L, M, N, dx, x, y, z):
    """GR pair histogram of the CVD atom fractions (2 per pair)."""
    lib = load()
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    hist = np.zeros(lib.pfo_cvd_gr_bins(L, ctypes.c_float(dx)), np.uint32)
    lib.pfo_cvd_gr(L, M, N, ctypes.c_float(dx), x.size, _fp(x), _fp(y), _fp(z),
                   hist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    return hist


def cvd_q6(L, M, N, dx, x, y, z):
    """Q6_GLO of the CVD atom fractions: (neighbour counts, local Q6,
    global Q6, 30-bin histogram of the local values)."""
    lib = load()
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    n = x.size
    nb = np.zeros(n, np.uint32)
    ql = np.zeros(n, np.float32)
    qg = np.zeros(1, np.float32)
    hq = np.zeros(30, np.uint64)
    lib.pfo_cvd_q6(L, M, N, ctypes.c_float(dx), n, _fp(x), _fp(y), _fp(z),
                   nb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), _fp(ql), _fp(qg),
                   hq.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)))
    return nb, ql, float(qg[0]), hq


def ylm(l, m, theta, fai):
    """(real, imaginary) spherical harmonic as the reference evaluates it (float)."""
    lib = load()
    return lib.pfo_ylm_r(l, m, theta, fai), lib.pfo_ylm_c(l, m, theta, fai)


def fft3d(z, L, M, N, sign):
    lib = load()
    a = np.ascontiguousarray(z, np.complex128).copy()
    lib.pfo_fft3d(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), L, M, N, sign)
    return a


def gauss(seed, timestep, n):
    """The noise stream of one step (pfo_gauss): (values, uniform pairs drawn)."""
    out = np.zeros(n, np.float64)
    pairs = load().pfo_gauss(seed, timestep, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out, int(pairs)
