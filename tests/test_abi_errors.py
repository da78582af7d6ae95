"""Error behaviour of the C ABI (include/pfc_hip.h) on paths that end before
any GPU call: every entry point returns a negative PFC_E_* code for bad
arguments and leaves a message in pfc_last_error(); nothing is allocated or
launched.  Also the GR bin-count query, which needs no GPU."""
import ctypes

import numpy as np
import pytest

E_ARG, E_UNSUPPORTED = -1, -2


def _err(lib):
    return lib.pfc_last_error().decode()


@pytest.mark.parametrize("field,value,code,msg", [
    ("model", 7, E_ARG, "bad model"),
    ("precision", 3, E_ARG, "bad precision"),
    ("L", 100, E_UNSUPPORTED, "not supported"),
    ("M", 2048, E_UNSUPPORTED, "not supported"),
    ("N", 8, E_UNSUPPORTED, "N = 8"),
    ("dt", 0.0, E_ARG, "dt and dx"),
    ("dx", -1.0, E_ARG, "dt and dx"),
])
def test_create_rejects_bad_params(hip, field, value, code, msg):
    lib = hip.load()
    p = hip.default_params(hip.MISFIT, L=64, M=64)
    setattr(p, field, value)
    h = ctypes.c_void_p(123)
    assert lib.pfc_create(ctypes.byref(p), 1, ctypes.byref(h)) == code
    assert msg in _err(lib)
    assert not h.value  # *out cleared before validation ends


def test_null_and_range_arguments(hip):
    lib = hip.load()
    p = hip.default_params(hip.MISFIT, L=64, M=64)
    h = ctypes.c_void_p()
    assert lib.pfc_create(None, 1, ctypes.byref(h)) == E_ARG
    assert lib.pfc_create(ctypes.byref(p), 0, ctypes.byref(h)) == E_ARG
    assert lib.pfc_step(None, 1) == E_ARG
    assert lib.pfc_set_graph(None, 4) == E_ARG
    assert lib.pfc_get_fields(None, 0, None, None) == E_ARG
    assert lib.pfc_set_col_tiles(None, 0) == E_ARG
    assert lib.pfc_destroy(None) == 0  # destroying nothing is fine
    assert lib.pfc_group_create(None, 0, 25, ctypes.byref(h)) == E_ARG
    assert lib.pfc_group_step(None, 1) == E_ARG
    phi = np.zeros(16, np.float64)
    bad = hip.default_params(hip.MISFIT, L=4, M=4)
    bad.grainx = 0
    assert lib.pfc_init_fields(ctypes.byref(bad), hip._dp(phi), hip._dp(phi)) == E_ARG
    assert "grain" in _err(lib)
    n = ctypes.c_int()
    assert lib.pfc_phimax_elastic(0, 64, 1, ctypes.c_float(0.5), None, None, None, None, None, 0,
                                  ctypes.byref(n)) == E_ARG
    assert lib.pfc_strain(64, 64, 1, ctypes.c_float(0.5), ctypes.c_float(1.7), 5, None, None, None, None,
                          None) == E_ARG
    assert lib.pfc_cryst_order(64, 64, 1, ctypes.c_float(0.5), -1, None, None, None, None, None, None) == E_ARG
    assert lib.pfc_cvd_q6(64, 64, 1, ctypes.c_float(0.5), 0, None, None, None, None, None, None, None) == E_ARG
    assert lib.pfc_copy_probe(16, 0, None) == E_ARG
    assert "bad arguments" in _err(lib)


def test_empty_inputs_are_no_ops(hip):
    lib = hip.load()
    assert lib.pfc_strain(64, 64, 1, ctypes.c_float(0.5), ctypes.c_float(1.7), 0, None, None, None, None,
                          None) == 0
    assert lib.pfc_cryst_order(64, 64, 1, ctypes.c_float(0.5), 0, None, None, None, None, None, None) == 0


@pytest.mark.parametrize("L,dx", [(64, 0.5), (1024, 0.5), (256, 0.125), (384, 0.3)])
def test_cvd_gr_bin_count_query(hip, L, dx):
    """With no histogram buffer pfc_cvd_gr only reports the bin count: the
    reference's maxbin = int(L*dx / delr) in float, delr = float(1/25) (cvd.cpp:1278-1279)."""
    lib = hip.load()
    nbin = ctypes.c_int(-1)
    assert lib.pfc_cvd_gr(L, L, 1, ctypes.c_float(dx), 0, None, None, None, None, ctypes.byref(nbin)) == 0
    delr = np.float32(1.0 / 25.0)
    assert nbin.value == int(np.float32(np.float32(L) * np.float32(dx)) / delr)
    assert lib.pfc_cvd_gr(L, L, 1, ctypes.c_float(0.0), 0, None, None, None, None, ctypes.byref(nbin)) == E_ARG
