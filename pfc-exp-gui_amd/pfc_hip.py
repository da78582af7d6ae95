"""ctypes binding of libpfc_hip.so (the C ABI declared in include/pfc_hip.h).

Host-side mirror of the reference's per-run interface: a context is one
`pfc-exp-cli --model misfit|cvd` run (or a batch of them), `step` is the body
of the reference's `for t` loop (misfit.cpp:477-498), `energy` is ENERGYC2,
`phimax_*` / `strain` are the checkpoint extractors.  Loading fails loudly
when the library has not been built: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfc_hip.so")

MISFIT, CVD = 0, 1
FP64, FP32 = 0, 1
KID_ROW, KID_COL, KID_DEPTH = 0, 1, 2

# Exported symbols (checked by the CPU test suite against include/pfc_hip.h).
SYMBOLS = [
    "pfc_params_default", "pfc_init_fields", "pfc_create", "pfc_destroy", "pfc_set_fields", "pfc_set_seed",
    "pfc_set_active",
    "pfc_step", "pfc_sync", "pfc_get_fields", "pfc_get_fields_f32", "pfc_energy",
    "pfc_step_energy", "pfc_step_energy_enqueue", "pfc_step_energy_result", "pfc_query", "pfc_phimax_misfit", "pfc_phimax_cvd", "pfc_phimax_elastic", "pfc_phimax_elastic_noalloc",
    "pfc_strain", "pfc_cryst_order", "pfc_cvd_gr", "pfc_cvd_q6", "pfc_col_tiles", "pfc_set_col_tiles", "pfc_group_create", "pfc_group_step",
    "pfc_group_destroy", "pfc_set_graph", "pfc_time_kernel", "pfc_kernel_time", "pfc_kernel_bytes",
    "pfc_device_memory", "pfc_copy_probe", "pfc_last_error",
    "pfc_server_create", "pfc_server_submit", "pfc_server_poll", "pfc_server_result", "pfc_server_error",
    "pfc_server_cancel", "pfc_server_drain", "pfc_server_stats", "pfc_server_destroy",
]

JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_CANCELLED, JOB_FAILED = range(5)


class Params(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int), ("L", ctypes.c_int), ("M", ctypes.c_int), ("N", ctypes.c_int),
        ("precision", ctypes.c_int), ("narrow", ctypes.c_int),
        ("u0", ctypes.c_float), ("con0", ctypes.c_float), ("sig", ctypes.c_float),
        ("dt", ctypes.c_float), ("dx", ctypes.c_float), ("axx", ctypes.c_float),
        ("grainx", ctypes.c_int), ("grainy", ctypes.c_int), ("grainz", ctypes.c_int),
        ("seed", ctypes.c_uint),
    ]


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_double), ("max_slots", ctypes.c_int), ("graph_chunk", ctypes.c_int),
                ("quantum", ctypes.c_int)]


class Job(ctypes.Structure):
    _fields_ = [("p", Params), ("steps", ctypes.c_int), ("mod", ctypes.c_int), ("keep_fields", ctypes.c_int),
                ("outdir", ctypes.c_char_p)]


class JobStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("steps_done", ctypes.c_int), ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_end", ctypes.c_double), ("energy", ctypes.c_double),
                ("density", ctypes.c_double), ("t_stepped", ctypes.c_double), ("create_s", ctypes.c_double)]


class ServerInfo(ctypes.Structure):
    _fields_ = [("now", ctypes.c_double), ("jobs_done", ctypes.c_longlong), ("jobs_unfinished", ctypes.c_int),
                ("point_steps", ctypes.c_double), ("contexts", ctypes.c_int), ("bytes_in_use", ctypes.c_double),
                ("budget", ctypes.c_double)]


class PfcError(RuntimeError):
    pass


_lib = None


def load():
    """Load libpfc_hip.so; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PfcError(f"{LIB_PATH} not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    sig = {
        "pfc_params_default": [ctypes.POINTER(Params), ctypes.c_int],
        "pfc_init_fields": [ctypes.POINTER(Params), dp, dp],
        "pfc_create": [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(vp)],
        "pfc_destroy": [vp],
        "pfc_set_fields": [vp, ctypes.c_int, dp, dp],
        "pfc_step": [vp, ctypes.c_int],
        "pfc_set_seed": [vp, ctypes.c_int, ctypes.c_uint],
        "pfc_set_active": [vp, ctypes.c_int, ctypes.c_int],
        "pfc_sync": [vp],
        "pfc_get_fields": [vp, ctypes.c_int, dp, dp],
        "pfc_get_fields_f32": [vp, ctypes.c_int, fp, fp],
        "pfc_energy": [vp, dp],
        "pfc_step_energy": [vp, ctypes.c_int, dp],
        "pfc_step_energy_enqueue": [vp, ctypes.c_int],
        "pfc_step_energy_result": [vp, ctypes.c_int, dp],
        "pfc_query": [vp, ctypes.POINTER(ctypes.c_int)],
        "pfc_phimax_misfit": [vp, ctypes.c_int, dp, dp, dp, dp, fp, ctypes.c_int, ip],
        "pfc_phimax_cvd": [vp, ctypes.c_int, fp, fp, fp, fp, ctypes.c_int, ip, fp, fp],
        "pfc_phimax_elastic": [ctypes.c_int] * 3 + [ctypes.c_float, fp, dp, dp, dp, dp, ctypes.c_int, ip],
        "pfc_phimax_elastic_noalloc": [ctypes.c_int] * 3 + [ctypes.c_float, vp, vp, vp, vp, vp, ctypes.c_int, vp, vp],
        "pfc_strain": [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_float, ctypes.c_int, dp, dp, dp, dp,
                                            ctypes.POINTER(ctypes.c_uint)],
        "pfc_cryst_order": [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_int, dp, dp, dp, dp, dp,
                                                 ctypes.POINTER(ctypes.c_uint)],
        "pfc_cvd_gr": [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_int, fp, fp, fp,
                                            ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_int)],
        "pfc_cvd_q6": [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_int, fp, fp, fp,
                                            ctypes.POINTER(ctypes.c_uint), fp, fp,
                                            ctypes.POINTER(ctypes.c_ulonglong)],
        "pfc_col_tiles": [vp, ctypes.POINTER(ctypes.c_int)],
        "pfc_set_col_tiles": [vp, ctypes.c_int],
        "pfc_group_create": [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)],
        "pfc_group_step": [vp, ctypes.c_int],
        "pfc_group_destroy": [vp],
        "pfc_set_graph": [vp, ctypes.c_int],
        "pfc_time_kernel": [vp, ctypes.c_int, ctypes.c_int],
        "pfc_kernel_time": [vp, dp, ip],
        "pfc_kernel_bytes": [vp, ctypes.c_int, dp],
        "pfc_device_memory": [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
        "pfc_copy_probe": [ctypes.c_size_t, ctypes.c_int, dp],
        "pfc_server_create": [ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)],
        "pfc_server_submit": [vp, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_longlong)],
        "pfc_server_poll": [vp, ctypes.c_longlong, ctypes.POINTER(JobStatus)],
        "pfc_server_result": [vp, ctypes.c_longlong, dp, dp],
        "pfc_server_error": [vp, ctypes.c_longlong, ctypes.c_char_p, ctypes.c_int],
        "pfc_server_cancel": [vp, ctypes.c_longlong],
        "pfc_server_drain": [vp, ctypes.c_double],
        "pfc_server_stats": [vp, ctypes.POINTER(ServerInfo)],
        "pfc_server_destroy": [vp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    lib.pfc_last_error.argtypes = []
    lib.pfc_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise PfcError(f"pfc error {rc}: {_lib.pfc_last_error().decode()}")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def copy_probe(nbytes, pattern=0):
    """Milliseconds of one plain device copy of nbytes (counter calibration);
    pattern 0 linear, 1 column-wise 64-byte pieces, 2 the same in the column
    pass's XCD-aware tile order."""
    ms = ctypes.c_double()
    _check(load().pfc_copy_probe(nbytes, pattern, ctypes.byref(ms)))
    return ms.value


def default_params(model=MISFIT, L=256, M=256, N=1, precision=FP64, **kw):
    lib = load()
    p = Params()
    _check(lib.pfc_params_default(ctypes.byref(p), model))
    p.L, p.M, p.N, p.precision = L, M, N, precision
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def init_fields(p):
    lib = load()
    P = p.L * p.M * p.N
    phi = np.zeros(P, np.float64)
    con = np.zeros(P, np.float64)
    _check(lib.pfc_init_fields(ctypes.byref(p), _dp(phi), _dp(con)))
    return phi, con


class Sim:
    """One batch of PFC instances on the GPU."""

    def __init__(self, p, batch=1):
        self.lib = load()
        self.p = p
        self.batch = batch
        self.P = p.L * p.M * p.N
        h = ctypes.c_void_p()
        _check(self.lib.pfc_create(ctypes.byref(p), batch, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.pfc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, phi, con, inst=0):
        phi = np.ascontiguousarray(phi, np.float64).ravel()
        con = np.ascontiguousarray(con, np.float64).ravel()
        assert phi.size == self.P and con.size == self.P
        _check(self.lib.pfc_set_fields(self.h, inst, _dp(phi), _dp(con)))

    def set_seed(self, seed, inst=0):
        _check(self.lib.pfc_set_seed(self.h, inst, seed))

    def set_active(self, inst, active=True):
        _check(self.lib.pfc_set_active(self.h, inst, 1 if active else 0))

    def step(self, n=1):
        _check(self.lib.pfc_step(self.h, n))

    def sync(self):
        _check(self.lib.pfc_sync(self.h))

    def get_fields(self, inst=0):
        phi = np.zeros(self.P, np.float64)
        con = np.zeros(self.P, np.float64)
        _check(self.lib.pfc_get_fields(self.h, inst, _dp(phi), _dp(con)))
        return phi, con

    def energy(self):
        e = np.zeros(2 * self.batch, np.float64)
        _check(self.lib.pfc_energy(self.h, _dp(e)))
        return e.reshape(self.batch, 2)

    def step_energy(self, n):
        tr = np.zeros(2 * self.batch * n, np.float64)
        _check(self.lib.pfc_step_energy(self.h, n, _dp(tr)))
        return tr.reshape(n, self.batch, 2)

    def set_graph(self, chunk):
        _check(self.lib.pfc_set_graph(self.h, chunk))

    def col_tiles(self):
        """Columns per workgroup the autotuner chose for the plain step."""
        t = ctypes.c_int()
        _check(self.lib.pfc_col_tiles(self.h, ctypes.byref(t)))
        return t.value

    def set_col_tiles(self, candidate):
        """Force tile-width candidate 0 (default), 1 (twice) or 2 (half)."""
        _check(self.lib.pfc_set_col_tiles(self.h, candidate))

    def time_kernel(self, kid, enable=True):
        _check(self.lib.pfc_time_kernel(self.h, kid, 1 if enable else 0))

    def kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        _check(self.lib.pfc_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def kernel_bytes(self, kid):
        b = ctypes.c_double()
        _check(self.lib.pfc_kernel_bytes(self.h, kid, ctypes.byref(b)))
        return b.value

    def phimax_misfit(self, inst=0):
        cap = self.P
        x, y, z, ph = (np.zeros(cap, np.float64) for _ in range(4))
        co = np.zeros(cap, np.float32)
        n = ctypes.c_int()
        _check(self.lib.pfc_phimax_misfit(self.h, inst, _dp(x), _dp(y), _dp(z), _dp(ph), _fp(co), cap,
                                          ctypes.byref(n)))
        k = n.value
        return x[:k], y[:k], z[:k], ph[:k], co[:k]

    def phimax_cvd(self, inst=0):
        cap = self.P
        px, py, pz, pp = (np.zeros(cap, np.float32) for _ in range(4))
        n = ctypes.c_int()
        phic = ctypes.c_float()
        pmax = ctypes.c_float()
        _check(self.lib.pfc_phimax_cvd(self.h, inst, _fp(px), _fp(py), _fp(pz), _fp(pp), cap, ctypes.byref(n),
                                       ctypes.byref(phic), ctypes.byref(pmax)))
        k = n.value
        return px[:k], py[:k], pz[:k], pp[:k], phic.value, pmax.value


def phimax_elastic(L, M, N, dx, phi):
    lib = load()
    phi = np.ascontiguousarray(phi, np.float32).ravel()
    cap = phi.size
    x, y, z, ph = (np.zeros(cap, np.float64) for _ in range(4))
    n = ctypes.c_int()
    _check(lib.pfc_phimax_elastic(L, M, N, ctypes.c_float(dx), _fp(phi), _dp(x), _dp(y), _dp(z), _dp(ph), cap,
                                  ctypes.byref(n)))
    k = n.value
    return x[:k], y[:k], z[:k], ph[:k]


def strain(L, M, N, dx, benchel, x, y, z):
    lib = load()
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    n = x.size
    s = np.zeros(n, np.float64)
    nb = np.zeros(n, np.uint32)
    _check(lib.pfc_strain(L, M, N, ctypes.c_float(dx), ctypes.c_float(benchel), n, _dp(x), _dp(y), _dp(z), _dp(s),
                          nb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))))
    return s, nb


def cryst_order(L, M, N, dx, x, y, z):
    """Per-atom (q4, q6, neighbour count) of cryst_order."""
    lib = load()
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    n = x.size
    q4 = np.zeros(n, np.float64)
    q6 = np.zeros(n, np.float64)
    nb = np.zeros(n, np.uint32)
    _check(lib.pfc_cryst_order(L, M, N, ctypes.c_float(dx), n, _dp(x), _dp(y), _dp(z), _dp(q4), _dp(q6),
                               nb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))))
    return q4, q6, nb


def cvd_gr(L, M, N, dx, x, y, z):
    """GR pair histogram of CVD atom fractions (pfc_phimax_cvd's pc_x/y/z)."""
    lib = load()
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    nbin = ctypes.c_int(0)
    _check(lib.pfc_cvd_gr(L, M, N, ctypes.c_float(dx), x.size, _fp(x), _fp(y), _fp(z), None, ctypes.byref(nbin)))
    hist = np.zeros(nbin.value, np.uint32)
    _check(lib.pfc_cvd_gr(L, M, N, ctypes.c_float(dx), x.size, _fp(x), _fp(y), _fp(z),
                          hist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), ctypes.byref(nbin)))
    return hist


def cvd_q6(L, M, N, dx, x, y, z):
    """Q6_GLO of CVD atom fractions: (neighbour counts, local Q6, global Q6,
    30-bin histogram)."""
    lib = load()
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    n = x.size
    nb = np.zeros(n, np.uint32)
    ql = np.zeros(n, np.float32)
    qg = np.zeros(1, np.float32)
    hq = np.zeros(30, np.uint64)
    _check(lib.pfc_cvd_q6(L, M, N, ctypes.c_float(dx), n, _fp(x), _fp(y), _fp(z),
                          nb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), _fp(ql), _fp(qg),
                          hq.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong))))
    return nb, ql, float(qg[0]), hq


class Group:
    """Several Sims stepped as one graph (chunk steps of each per launch)."""

    def __init__(self, sims, chunk):
        self.lib = load()
        self.sims = list(sims)
        self.chunk = chunk
        arr = (ctypes.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        self.h = ctypes.c_void_p()
        _check(self.lib.pfc_group_create(arr, len(self.sims), chunk, ctypes.byref(self.h)))

    def step(self, nchunks=1):
        _check(self.lib.pfc_group_step(self.h, nchunks))

    def sync(self):
        for s in self.sims:
            s.sync()

    def close(self):
        if self.h:
            self.lib.pfc_group_destroy(self.h)
            self.h = ctypes.c_void_p()


class Server:
    """Request loop with continuous batching (pfc_server_* of the C ABI)."""

    def __init__(self, hbm_budget=0.0, max_slots=16, graph_chunk=25, quantum=500):
        self.lib = load()
        o = ServerOpts(hbm_budget, max_slots, graph_chunk, quantum)
        h = ctypes.c_void_p()
        _check(self.lib.pfc_server_create(ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, p, steps, mod=0, keep_fields=False, outdir=None):
        j = Job(p, steps, mod, 1 if keep_fields else 0, outdir.encode() if outdir else None)
        jid = ctypes.c_longlong()
        _check(self.lib.pfc_server_submit(self.h, ctypes.byref(j), ctypes.byref(jid)))
        self._keep[jid.value] = p.L * p.M * p.N
        return jid.value

    def poll(self, jid):
        st = JobStatus()
        _check(self.lib.pfc_server_poll(self.h, jid, ctypes.byref(st)))
        return st

    def result(self, jid):
        P = self._keep[jid]
        phi = np.zeros(P, np.float64)
        con = np.zeros(P, np.float64)
        _check(self.lib.pfc_server_result(self.h, jid, _dp(phi), _dp(con)))
        return phi, con

    def error(self, jid):
        buf = ctypes.create_string_buffer(512)
        _check(self.lib.pfc_server_error(self.h, jid, buf, 512))
        return buf.value.decode()

    def cancel(self, jid):
        _check(self.lib.pfc_server_cancel(self.h, jid))

    def drain(self, timeout=-1.0):
        _check(self.lib.pfc_server_drain(self.h, timeout))

    def stats(self):
        st = ServerInfo()
        _check(self.lib.pfc_server_stats(self.h, ctypes.byref(st)))
        return st

    def close(self):
        if self.h:
            self.lib.pfc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
