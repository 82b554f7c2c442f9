"""kw_spectral_gradient_noalloc against kw_test_spectral_gradient, through the C ABI.

The allocation-free entry point takes device arrays and generates its line
tables into the library's static scratch array; the allocating one uploads
host arrays and host-built tables.  Both run the same derivative kernel, so
FP64 results agree to rounding of the table generation (<= 1e-13 relative L2)
and FP32 results to FP32 table rounding (<= 1e-6).  Two calls in a row, the
second on a smaller grid, check that the scratch left by the first call does
not leak into the second.  The test's own device buffers come from the HIP
runtime the library links (allocating them is the caller's business).
"""
import ctypes as C

import numpy as np
import pytest

from kwave_amd import _lib

pytestmark = pytest.mark.gpu

_DP = C.POINTER(C.c_double)


def _hip():
    _lib.load()
    h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


H2D, D2H = 1, 2


def _check_hip(rc, what):
    assert rc == 0, f"{what} returned hipError {rc}"


def _reference(f, d, axis, sign, dtype):
    lib = _lib.load()
    nd = f.ndim
    N = (C.c_int64 * 3)(*(list(f.shape) + [1] * (3 - nd)))
    dd = (C.c_double * 3)(*(list(d) + [1.0] * (3 - nd)))
    src = np.ascontiguousarray(f.ravel(order="F"), dtype=np.float64)
    out = np.empty_like(src)
    _lib.check(lib.kw_test_spectral_gradient(_lib.context(0), 4 if dtype == np.float32 else 8, nd, N, dd, axis,
                                             sign, src.ctypes.data_as(_DP), out.ctypes.data_as(_DP)))
    return out


def _noalloc(hip, stream, f, d, axis, sign, dtype):
    lib = _lib.load()
    nd = f.ndim
    N = (C.c_int64 * 3)(*(list(f.shape) + [1] * (3 - nd)))
    dd = (C.c_double * 3)(*(list(d) + [1.0] * (3 - nd)))
    host = np.ascontiguousarray(f.ravel(order="F"), dtype=dtype)
    nbytes = host.nbytes
    dev_f, dev_o = C.c_void_p(), C.c_void_p()
    _check_hip(hip.hipMalloc(C.byref(dev_f), nbytes), "hipMalloc")
    _check_hip(hip.hipMalloc(C.byref(dev_o), nbytes), "hipMalloc")
    try:
        _check_hip(hip.hipMemcpy(dev_f, host.ctypes.data, nbytes, H2D), "hipMemcpy")
        _lib.check(lib.kw_spectral_gradient_noalloc(_lib.context(0), 4 if dtype == np.float32 else 8, nd, N, dd,
                                                    axis, sign, dev_f, dev_o, stream))
        _check_hip(hip.hipStreamSynchronize(stream), "hipStreamSynchronize")
        out = np.empty_like(host)
        _check_hip(hip.hipMemcpy(out.ctypes.data, dev_o, nbytes, D2H), "hipMemcpy")
    finally:
        hip.hipFree(dev_f)
        hip.hipFree(dev_o)
    return out.astype(np.float64)


CASES = [
    # (shape of the first call, shape of the second (smaller) call, spacing, axis, sign)
    ((256, 128, 40), (64, 48, 20), (1e-4, 2e-4, 1.5e-4), 1, +1),
    ((128, 96, 64), (40, 24, 16), (1e-4, 1e-4, 1e-4), 2, -1),
    ((512, 64), (96, 45), (1e-4, 1e-4), 0, 0),
    ((256, 30), (64, 9), (1e-4, 1e-4), 0, +1),
    ((243,), (45,), (1e-4,), 0, -1),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-axis{c[3]}-sign{c[4]}" for c in CASES])
def test_noalloc_matches_allocating_entry_point(case, dtype):
    big, small, d, axis, sign = case
    print(f"\n[case] noalloc derivative {big} then {small}, axis {axis}, sign {sign}, {dtype.__name__}", flush=True)
    hip = _hip()
    stream = C.c_void_p()
    _check_hip(hip.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    tol = 1e-13 if dtype == np.float64 else 1e-6
    rng = np.random.default_rng(7)
    try:
        for shape in (big, small):
            f = rng.standard_normal(shape)
            ref = _reference(f, d, axis, sign, dtype)
            got = _noalloc(hip, stream, f, d, axis, sign, dtype)
            e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            print(f"    {shape}: rel-L2 {e:.3e}", flush=True)
            assert np.all(np.isfinite(got))
            assert e <= tol, (shape, e)
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_rejects_bad_arguments():
    lib = _lib.load()
    N = (C.c_int64 * 3)(7, 8, 1)
    dd = (C.c_double * 3)(1e-4, 1e-4, 1.0)
    one = C.c_void_p(1)     # never dereferenced: the checks fail before anything is enqueued
    rc = lib.kw_spectral_gradient_noalloc(_lib.context(0), 8, 2, N, dd, 0, 1, one, one, None)
    assert rc == _lib.KW_ERR_UNSUPPORTED            # 7 has no plan
    N = (C.c_int64 * 3)(8, 8, 1)
    rc = lib.kw_spectral_gradient_noalloc(_lib.context(0), 2, 2, N, dd, 0, 1, one, one, None)
    assert rc == _lib.KW_ERR_INVALID                # precision
    rc = lib.kw_spectral_gradient_noalloc(_lib.context(0), 8, 2, N, dd, 1, 0, one, one, None)
    assert rc == _lib.KW_ERR_UNSUPPORTED            # unstaggered y derivative of an even axis
    rc = lib.kw_spectral_gradient_noalloc(_lib.context(0), 8, 2, N, dd, 0, 1, None, one, None)
    assert rc == _lib.KW_ERR_INVALID                # null input
