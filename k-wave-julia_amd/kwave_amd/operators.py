"""Single operators of the solvers through the C ABI (kw_test_*): the host
tables and one spectral derivative on the GPU.  For operator-level parity
tests against the reference's formulas (grid.jl, pml.jl, fft_utils.jl,
first_order_steps.jl)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .solver import _precision

_DP = C.POINTER(C.c_double)


def axis_table(n: int, d: float, which: str = "k", pml_size: int = 0, pml_alpha: float = 0.0,
               dt: float = 0.0) -> np.ndarray:
    """k vector, PML profiles or staggered derivative multipliers of one axis, as the solvers build them."""
    lib = _lib.load()
    w = _lib.AXIS_TABLE[which]
    cplx = which in ("dplus", "dminus")
    out = np.empty(2 * n if cplx else n, dtype=np.float64)
    _lib.check(lib.kw_test_axis_table(int(n), float(d), int(pml_size), float(pml_alpha), float(dt), w,
                                      out.ctypes.data_as(_DP)))
    return out[0::2] + 1j * out[1::2] if cplx else out


def eos_table(desc: "_lib.SimDesc", which: int = 0) -> np.ndarray:
    """EOS multiplier on the rfft grid (Nx//2+1, Ny, Nz) for a simulation descriptor."""
    lib = _lib.load()
    nd = desc.ndim
    shape = [desc.N[0] // 2 + 1] + [desc.N[a] for a in range(1, nd)]
    out = np.empty(int(np.prod(shape)), dtype=np.float64)
    _lib.check(lib.kw_test_eos_table(C.byref(desc), int(which), out.ctypes.data_as(_DP)))
    return out.reshape(shape, order="F")


def spectral_gradient(f: np.ndarray, d, axis: int, sign: int = 1, data_cast=np.float64, device: int = 0) -> np.ndarray:
    """d f / d x_axis by one GPU line pass: sign +1 / -1 staggered (e^{+-i k d/2}), 0 unstaggered."""
    lib = _lib.load()
    f = np.asarray(f, dtype=np.float64)
    nd = f.ndim
    N = (C.c_int64 * 3)(*(list(f.shape) + [1] * (3 - nd)))
    dd = (C.c_double * 3)(*(list(np.broadcast_to(np.asarray(d, dtype=np.float64), (nd,))) + [1.0] * (3 - nd)))
    src = np.ascontiguousarray(f.ravel(order="F"))
    out = np.empty_like(src)
    _lib.check(lib.kw_test_spectral_gradient(_lib.context(device), _precision(data_cast), nd, N, dd, int(axis),
                                             int(sign), src.ctypes.data_as(_DP), out.ctypes.data_as(_DP)))
    return out.reshape(f.shape, order="F")
