"""pfc_phimax_elastic_noalloc (the allocation-free elastic Phimax: device
field in, device atoms out, scratch in a fixed device array) against
pfc_phimax_elastic (host arrays, library-allocated scratch) on the same
fields, through the C ABI: two calls in a row, the second on a smaller field
(so stale scratch from the first would show), then a capacity below the atom
count and a field beyond the scratch's maximum size."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Dev:
    """Device buffers through the HIP runtime (the caller's memory)."""

    def __init__(self):
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.h = h
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(nbytes, 16)) == 0
        self.bufs.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype)
        if n:
            assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _field(hip, model, L, M, N, steps):
    kw = dict(grainx=max(1, L // 32), grainy=max(1, M // 32)) if model == hip.MISFIT else {}
    p = hip.default_params(model=model, L=L, M=M, N=N, precision=hip.FP64, **kw)
    phi, con = hip.init_fields(p)
    s = hip.Sim(p)
    s.set_fields(phi, con)
    s.step(steps)
    f = s.get_fields()[0].astype(np.float32)
    s.close()
    return f


def _noalloc(hip, dev, stream, L, M, N, dphi, cap):
    lib = hip.load()
    outs = [dev.alloc(8 * cap) for _ in range(4)]
    dn = dev.alloc(4)
    rc = lib.pfc_phimax_elastic_noalloc(L, M, N, ctypes.c_float(0.125), dphi, *outs, cap, dn, stream)
    assert rc == 0, lib.pfc_last_error().decode()
    assert dev.h.hipStreamSynchronize(stream) == 0
    n = int(dev.get(dn, 1, np.int32)[0])
    k = min(n, cap)
    return n, [dev.get(o, k, np.float64) for o in outs]


def test_noalloc_matches_host_entry_point(hip):
    dev = Dev()
    stream = ctypes.c_void_p()
    assert dev.h.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # a 256^2 Misfit field, then a smaller 3-D CVD field (32x32x16)
        cases = [(hip.MISFIT, 256, 256, 1, 60), (hip.CVD, 32, 32, 16, 5)]
        for model, L, M, N, steps in cases:
            phi = _field(hip, model, L, M, N, steps)
            print(f"noalloc case model={model} grid={L}x{M}x{N}", flush=True)
            hx, hy, hz, hp = hip.phimax_elastic(L, M, N, 0.125, phi)
            assert len(hx) > 10
            dphi = dev.put(phi)
            n, (x, y, z, ph) = _noalloc(hip, dev, stream, L, M, N, dphi, phi.size)
            assert n == len(hx)
            for a, b in ((x, hx), (y, hy), (z, hz), (ph, hp)):
                assert np.array_equal(a, b)
            # capacity below the count: the full count, the first cap atoms
            cap = n // 3
            n2, (x2, y2, z2, p2) = _noalloc(hip, dev, stream, L, M, N, dphi, cap)
            assert n2 == n and len(x2) == cap
            assert np.array_equal(x2, hx[:cap]) and np.array_equal(p2, hp[:cap])
        # beyond the scratch's maximum: rejected before any GPU work
        lib = hip.load()
        big = dev.alloc(16)
        rc = lib.pfc_phimax_elastic_noalloc(4096, 4096, 2, ctypes.c_float(0.125), big, big, big, big, big, 0, big,
                                            stream)
        assert rc == -2 and b"at most" in lib.pfc_last_error()
    finally:
        dev.h.hipStreamSynchronize(stream)
        dev.h.hipStreamDestroy(stream)
        dev.free()
