"""pn_volume_analysis_noalloc (the entry point that allocates nothing, for
callers under HIP graph capture or in serving loops) against
pn_volume_analysis on the same fields: the same kernels and merge order, so
the ten averages and the extrema are bitwise equal.  The fields live in device
buffers the test allocates itself, as such a caller's would; two calls in a
row on the shared static scratch, the second on a smaller lattice."""
import ctypes as C

import numpy as np
import pytest

import phinum_amd as pn

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    pn.lib()  # brings in the HIP runtime the library itself links
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _fields(N, seed):
    rng = np.random.default_rng(seed)
    phi = 2.4 + 0.6 * rng.standard_normal(N ** 3)
    pi = 0.3 * rng.standard_normal(N ** 3)
    # sites outside the fast-division range take the IEEE path in their wave
    phi[5] = 0.0
    phi[N ** 3 // 2] = -3e-310
    phi[-7] = 1e-40
    return phi, pi


def _via_context(N, prec, phi, pi):
    with pn.Lattice(N, 0.5, 1, 1.0, -1, 1.224744871391589, precision=prec) as L:
        L.set_fields(phi, pi)
        f, p = L.get_fields()
        avg, ext = L.volume_analysis()
    want = np.concatenate([avg.as_array(), [ext.min, ext.minpos, ext.max, ext.maxpos, ext.maxabs,
                                            ext.maxabspos]])
    return f, p, want


def _via_noalloc(hip, N, prec, f, p):
    dt = np.float64 if prec == pn.PN_FP64 else np.float32
    hf = np.ascontiguousarray(f.astype(dt))
    hp = np.ascontiguousarray(p.astype(dt))
    dphi, dpi, dout = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dphi), hf.nbytes) == 0
    assert hip.hipMalloc(C.byref(dpi), hp.nbytes) == 0
    assert hip.hipMalloc(C.byref(dout), pn.PN_VOLUME_OUT * 8) == 0
    try:
        assert hip.hipMemcpy(dphi, hf.ctypes.data, hf.nbytes, H2D) == 0
        assert hip.hipMemcpy(dpi, hp.ctypes.data, hp.nbytes, H2D) == 0
        params = pn.PnParams(N, 0.5, 1, 1.0, -1, 1.224744871391589, prec, 0)
        rc = pn.lib().pn_volume_analysis_noalloc(C.byref(params), dphi, dpi, None, dout)
        assert rc == pn.PN_OK, pn.lib().pn_last_error(None)
        assert hip.hipDeviceSynchronize() == 0
        out = np.empty(pn.PN_VOLUME_OUT)
        assert hip.hipMemcpy(out.ctypes.data, dout, out.nbytes, D2H) == 0
        return out
    finally:
        for d in (dphi, dpi, dout):
            hip.hipFree(d)


@pytest.mark.parametrize("prec,sizes", [(pn.PN_FP64, (64, 20)), (pn.PN_FP64, (33, 9)),
                                        (pn.PN_FP32, (48, 12))])
def test_noalloc_equals_context_entry_point(prec, sizes):
    hip = _hip()
    for k, N in enumerate(sizes):  # the second call reuses the scratch on a smaller lattice
        print("noalloc case: N=%d precision=%d call %d" % (N, prec, k), flush=True)
        phi, pi = _fields(N, 100 + N)
        f, p, want = _via_context(N, prec, phi, pi)
        got = _via_noalloc(hip, N, prec, f, p)
        assert np.array_equal(got, want), (N, got, want)


def test_noalloc_rejects_sizes_beyond_the_scratch():
    for N in (3, pn.PN_NOALLOC_MAX_N + 1):
        params = pn.PnParams(N, 0.5, 1, 1.0, -1, 2.0, pn.PN_FP64, 0)
        rc = pn.lib().pn_volume_analysis_noalloc(C.byref(params), C.c_void_p(8), C.c_void_p(8), None,
                                                 C.c_void_p(8))
        assert rc == pn.PN_ERR_ARG
        assert b"outside [4, 2048]" in pn.lib().pn_last_error(None)
