"""ctypes binding of libkwave_hip.so (include/kwave_hip.h).

The library is loaded from this package folder.  There is no fallback: if
the shared object is missing or no GPU is visible, calls raise
``KWaveHIPError`` -- the product path never computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkwave_hip.so")

KW_OK = 0
KW_ERR_INVALID = -1
KW_ERR_UNSUPPORTED = -2
KW_ERR_HIP = -3
KW_ERR_OOM = -4
KW_ERR_STATE = -5

REC_BITS = {
    "p": 1 << 0, "p_max": 1 << 1, "p_min": 1 << 2, "p_rms": 1 << 3, "p_final": 1 << 4,
    "ux": 1 << 5, "uy": 1 << 6, "uz": 1 << 7, "u_max": 1 << 8, "u_rms": 1 << 9,
    "u_final": 1 << 10, "I_avg": 1 << 11, "I_max": 1 << 12,
}
FIELD = {"p": 0, "ux": 1, "uy": 2, "uz": 3, "rhox": 4, "rhoy": 5, "rhoz": 6, "rho_prev": 7}
ALPHA_MODE = {"no_absorption": 0, "no_dispersion": 1, "stokes": 2}

# every symbol include/kwave_hip.h declares
EXPORTS = (
    "kw_abi_version", "kw_last_error", "kw_desc_sizes", "kw_device_count", "kw_ctx_create", "kw_ctx_destroy",
    "kw_fft_length_supported", "kw_grid_supported", "kw_sim_create", "kw_sim_run", "kw_sim_synchronize",
    "kw_sim_steps_done", "kw_sim_get_field", "kw_sim_set_field", "kw_sim_sensor_shape", "kw_sim_get_sensor",
    "kw_sim_profile", "kw_sim_grad_c0", "kw_sim_create_firings", "kw_sim_destroy", "kw_autotune", "kw_tune_export", "kw_tune_import",
    "kw_tune_clear", "kw_batch_create", "kw_batch_run", "kw_batch_synchronize",
    "kw_batch_sensor_shape", "kw_batch_get_sensor", "kw_batch_get_field", "kw_batch_destroy",
    "kw_elastic_create", "kw_elastic_run", "kw_elastic_synchronize", "kw_elastic_steps_done",
    "kw_elastic_get_field", "kw_elastic_get_pressure", "kw_elastic_sensor_shape", "kw_elastic_get_sensor",
    "kw_elastic_profile", "kw_elastic_destroy", "kw_test_axis_table", "kw_test_eos_table",
    "kw_test_spectral_gradient", "kw_spectral_gradient_noalloc",
)
AXIS_TABLE = {"k": 0, "pml": 1, "pml_sg": 2, "dplus": 3, "dminus": 4}
EL_FIELD = {"vx": 0, "vy": 1, "vz": 2, "sxx": 3, "syy": 4, "szz": 5, "sxy": 6, "sxz": 7, "syz": 8}


class KWaveHIPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class SimDesc(C.Structure):
    """Mirror of kw_sim_desc (include/kwave_hip.h)."""
    _fields_ = [
        ("ndim", C.c_int32),
        ("N", C.c_int64 * 3),
        ("d", C.c_double * 3),
        ("dt", C.c_double),
        ("Nt", C.c_int64),
        ("precision", C.c_int32),
        ("map_storage", C.c_int32),
        ("pml_size", C.c_int32 * 3),
        ("pml_alpha", C.c_double * 3),
        ("c_ref", C.c_double),
        ("c0_offset", C.c_double),
        ("rho0_offset", C.c_double),
        ("c0", C.POINTER(C.c_double)),
        ("c0_scalar", C.c_double),
        ("rho0", C.POINTER(C.c_double)),
        ("rho0_scalar", C.c_double),
        ("nonlinear", C.c_int32),
        ("bona", C.POINTER(C.c_double)),
        ("bona_scalar", C.c_double),
        ("absorbing", C.c_int32),
        ("alpha_mode", C.c_int32),
        ("alpha_coeff", C.c_double),
        ("alpha_power", C.c_double),
        ("p0", C.POINTER(C.c_double)),
        ("smooth_p0", C.c_int32),
        ("n_psrc", C.c_int64),
        ("psrc_index", C.POINTER(C.c_int64)),
        ("psrc_signal", C.POINTER(C.c_double)),
        ("psrc_ncols", C.c_int64),
        ("p_mode", C.c_int32),
        ("n_usrc", C.c_int64),
        ("usrc_index", C.POINTER(C.c_int64)),
        ("usrc_signal", C.POINTER(C.c_double) * 3),
        ("usrc_ncols", C.c_int64 * 3),
        ("u_mode", C.c_int32),
        ("n_sensor", C.c_int64),
        ("sensor_index", C.POINTER(C.c_int64)),
        ("record", C.c_uint32),
        ("tr_data", C.POINTER(C.c_double)),
        ("directivity", C.c_int32),
        ("directivity_angle", C.POINTER(C.c_double)),
        ("directivity_angle_scalar", C.c_double),
        ("directivity_size", C.c_double),
        ("freq_response", C.c_int32),
        ("freq_center", C.c_double),
        ("freq_bandwidth", C.c_double),
        ("axisymmetric", C.c_int32),
    ]


class ElasticDesc(C.Structure):
    """Mirror of kw_elastic_desc (include/kwave_hip.h)."""
    _fields_ = [
        ("ndim", C.c_int32),
        ("N", C.c_int64 * 3),
        ("d", C.c_double * 3),
        ("dt", C.c_double),
        ("Nt", C.c_int64),
        ("precision", C.c_int32),
        ("pml_size", C.c_int32 * 3),
        ("pml_alpha", C.c_double * 3),
        ("cp", C.POINTER(C.c_double)),
        ("cp_scalar", C.c_double),
        ("cs", C.POINTER(C.c_double)),
        ("cs_scalar", C.c_double),
        ("rho0", C.POINTER(C.c_double)),
        ("rho0_scalar", C.c_double),
        ("chi_p", C.c_double),
        ("chi_s", C.c_double),
        ("p0", C.POINTER(C.c_double)),
        ("n_usrc", C.c_int64),
        ("usrc_index", C.POINTER(C.c_int64)),
        ("usrc_signal", C.POINTER(C.c_double) * 3),
        ("usrc_ncols", C.c_int64 * 3),
        ("u_mode", C.c_int32),
        ("n_ssrc", C.c_int64),
        ("ssrc_index", C.POINTER(C.c_int64)),
        ("ssrc_signal", C.POINTER(C.c_double) * 6),
        ("ssrc_ncols", C.c_int64 * 6),
        ("s_mode", C.c_int32),
        ("n_sensor", C.c_int64),
        ("sensor_index", C.POINTER(C.c_int64)),
        ("record", C.c_uint32),
    ]


_lib = None
_lock = threading.Lock()


def load():
    """Load the shared library once (raises if it is missing)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise KWaveHIPError(KW_ERR_STATE, f"{LIB_PATH} is missing: build it with "
                                "`python k-wave-julia_amd/csrc/build.py` (hipcc, gfx950)")
        lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            "kw_abi_version": ([], C.c_int32),
            "kw_last_error": ([], C.c_char_p),
            "kw_desc_sizes": ([C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int32),
            "kw_device_count": ([C.POINTER(C.c_int32)], C.c_int32),
            "kw_ctx_create": ([C.c_int32, C.POINTER(P)], C.c_int32),
            "kw_ctx_destroy": ([P], C.c_int32),
            "kw_fft_length_supported": ([C.c_int64], C.c_int32),
            "kw_grid_supported": ([C.c_int32, C.POINTER(C.c_int64), C.c_int32], C.c_int32),
            "kw_sim_create": ([P, C.POINTER(SimDesc), C.POINTER(P)], C.c_int32),
            "kw_sim_run": ([P, C.c_int64], C.c_int32),
            "kw_sim_synchronize": ([P], C.c_int32),
            "kw_sim_steps_done": ([P, C.POINTER(C.c_int64)], C.c_int32),
            "kw_sim_get_field": ([P, C.c_int32, C.POINTER(C.c_double)], C.c_int32),
            "kw_sim_set_field": ([P, C.c_int32, C.POINTER(C.c_double)], C.c_int32),
            "kw_sim_sensor_shape": ([P, C.c_uint32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int32),
            "kw_sim_get_sensor": ([P, C.c_uint32, C.POINTER(C.c_double)], C.c_int32),
            "kw_sim_profile": ([P, C.c_int64, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_char_p),
                                C.POINTER(C.c_double), C.POINTER(C.c_int64)], C.c_int32),
            "kw_sim_grad_c0": ([P, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)], C.c_int32),
            "kw_sim_create_firings": ([P, C.POINTER(SimDesc), C.c_int32, C.POINTER(P)], C.c_int32),
            "kw_sim_destroy": ([P], C.c_int32),
            "kw_autotune": ([P, C.POINTER(SimDesc), C.c_int32, C.c_int32], C.c_int32),
            "kw_tune_export": ([C.c_char_p, C.c_int64, C.POINTER(C.c_int64)], C.c_int32),
            "kw_tune_import": ([C.c_char_p], C.c_int32),
            "kw_tune_clear": ([], C.c_int32),
            "kw_batch_create": ([P, C.POINTER(SimDesc), C.c_int32, C.c_int32, C.POINTER(P)], C.c_int32),
            "kw_batch_run": ([P, C.c_int64], C.c_int32),
            "kw_batch_synchronize": ([P], C.c_int32),
            "kw_batch_sensor_shape": ([P, C.c_int32, C.c_uint32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
                                      C.c_int32),
            "kw_batch_get_sensor": ([P, C.c_int32, C.c_uint32, C.POINTER(C.c_double)], C.c_int32),
            "kw_batch_get_field": ([P, C.c_int32, C.c_int32, C.POINTER(C.c_double)], C.c_int32),
            "kw_batch_destroy": ([P], C.c_int32),
            "kw_elastic_create": ([P, C.POINTER(ElasticDesc), C.POINTER(P)], C.c_int32),
            "kw_elastic_run": ([P, C.c_int64], C.c_int32),
            "kw_elastic_synchronize": ([P], C.c_int32),
            "kw_elastic_steps_done": ([P, C.POINTER(C.c_int64)], C.c_int32),
            "kw_elastic_get_field": ([P, C.c_int32, C.POINTER(C.c_double)], C.c_int32),
            "kw_elastic_get_pressure": ([P, C.POINTER(C.c_double)], C.c_int32),
            "kw_elastic_sensor_shape": ([P, C.c_uint32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int32),
            "kw_elastic_get_sensor": ([P, C.c_uint32, C.POINTER(C.c_double)], C.c_int32),
            "kw_elastic_profile": ([P, C.c_int64, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_char_p),
                                    C.POINTER(C.c_double), C.POINTER(C.c_int64)], C.c_int32),
            "kw_elastic_destroy": ([P], C.c_int32),
            "kw_test_axis_table": ([C.c_int64, C.c_double, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                    C.POINTER(C.c_double)], C.c_int32),
            "kw_test_eos_table": ([C.POINTER(SimDesc), C.c_int32, C.POINTER(C.c_double)], C.c_int32),
            "kw_test_spectral_gradient": ([P, C.c_int32, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                           C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)],
                                          C.c_int32),
            "kw_spectral_gradient_noalloc": ([P, C.c_int32, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                              C.c_int32, C.c_int32, P, P, P], C.c_int32),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        if lib.kw_abi_version() != 1:
            raise KWaveHIPError(KW_ERR_STATE, "libkwave_hip ABI version mismatch")
        # a library built from another version of include/kwave_hip.h would misread the
        # descriptors' pointers: refuse it here instead of crashing inside kw_sim_create
        ns, ne = C.c_int64(), C.c_int64()
        lib.kw_desc_sizes(C.byref(ns), C.byref(ne))
        if (ns.value, ne.value) != (C.sizeof(SimDesc), C.sizeof(ElasticDesc)):
            raise KWaveHIPError(KW_ERR_STATE, f"{LIB_PATH} was built from a different include/kwave_hip.h "
                                f"(descriptor sizes {ns.value}/{ne.value} bytes, this binding "
                                f"{C.sizeof(SimDesc)}/{C.sizeof(ElasticDesc)}): rebuild it")
        _lib = lib
        return lib


def check(code: int) -> None:
    if code != KW_OK:
        msg = _lib.kw_last_error().decode("utf-8", "replace") if _lib is not None else ""
        raise KWaveHIPError(code, msg or f"libkwave_hip error {code}")


_ctx = {}


def context(device: int = 0):
    """Process-wide device context for `device` (created on first use)."""
    lib = load()
    with _lock:
        if device in _ctx:
            return _ctx[device]
        n = C.c_int32(0)
        check(lib.kw_device_count(C.byref(n)))
        if n.value == 0:
            raise KWaveHIPError(KW_ERR_HIP, "no HIP device visible")
        h = C.c_void_p()
        check(lib.kw_ctx_create(int(device), C.byref(h)))
        _ctx[device] = h
        return h
