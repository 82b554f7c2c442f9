"""pstd_elastic_2d / pstd_elastic_3d: drop-ins for KWave.jl's elastic solvers.

Mirrors ``pstd_elastic_2d(kgrid, medium, source, sensor; pml_size=20,
pml_alpha=2.0, data_cast=Float64, plot_sim=false, progress_callback=nothing)``
(KWave.jl/src/solver/elastic.jl:133-143) and ``pstd_elastic_3d`` (same
keywords, pml_size=10; elastic.jl:386-396): same validation errors and
``SimulationOutput`` keys.  The host computes the Kelvin-Voigt factors and
index lists; the time loop, recording and finalisation run in libkwave_hip
(kw_elastic_*) on the GPU.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _lib
from .model import (ElasticMedium, ElasticSource, KWaveGrid2D, KWaveGrid3D, KWaveSensor, SimulationOutput,
                    SourceMode, validate_record_fields)
from .solver import _colmajor, _dptr, _findall, _iptr, _norm3, _pml_alpha_cast, _precision

_DP = C.POINTER(C.c_double)


def _db2neper(alpha: float, y: float) -> float:
    """conversion.jl:34 -- dB/(MHz^y cm) to Np/((rad/s)^y m)."""
    return alpha * (1e-2 / (20.0 * math.log10(math.e))) * (2.0 * math.pi * 1e6) ** (-y)


def _param(v, name: str):
    """(map or None, scalar) of a medium parameter (Float64 scalar or array)."""
    a = np.asarray(v, dtype=np.float64)
    if a.ndim == 0:
        return None, float(a)
    return _colmajor(a), 0.0


def _kelvin_voigt(medium: ElasticMedium, scalar_lame: bool):
    """elastic.jl:193-210: chi = 2 alpha_Np c_ref^(y-1) with c_ref the (mean) speed;
    the array branch of the stress update carries no viscous terms (elastic.jl:323-334)."""
    y = medium.alpha_power
    chi = [0.0, 0.0]
    for k, (coef, speed) in enumerate(((medium.alpha_coeff_compression, medium.sound_speed_compression),
                                       (medium.alpha_coeff_shear, medium.sound_speed_shear))):
        if coef is None or y is None:
            continue
        if np.ndim(coef) != 0:
            raise NotImplementedError("per-voxel elastic absorption coefficients are not part of this build")
        c_ref = float(speed) if np.ndim(speed) == 0 else float(np.mean(speed))
        chi[k] = 2.0 * _db2neper(float(coef), float(y)) * c_ref ** (float(y) - 1.0)
    return (chi[0], chi[1]) if scalar_lame else (0.0, 0.0)


def _signal_rows(sig, n: int, name: str) -> np.ndarray:
    """A source signal as an (n, ncols) Float64 matrix; a vector is one column
    (size(sig, 2) == 1 in elastic.jl:250-271)."""
    s = np.asarray(sig, dtype=np.float64)
    if s.ndim == 1:
        s = s[:, None]
    if s.ndim != 2:
        raise ValueError(f"source.{name} must be a vector or a matrix")
    if s.shape[0] < n:
        raise IndexError(f"source.{name} has {s.shape[0]} rows for {n} source points (BoundsError)")
    return np.asfortranarray(s[:n])


def build_elastic_desc(kgrid, medium: ElasticMedium, source: ElasticSource, sensor: KWaveSensor, *, pml_size,
                       pml_alpha, data_cast):
    nd = kgrid.ndim
    if kgrid.Nt == 0 or kgrid.dt == 0:
        raise ValueError("Time array not set. Call make_time!(kgrid, sound_speed) before running.")
    validate_record_fields(sensor)
    shape = tuple(kgrid.grid_size())
    keep: List[object] = []
    d = _lib.ElasticDesc()
    d.ndim = nd
    for a in range(3):
        d.N[a] = shape[a] if a < nd else 1
        d.d[a] = (kgrid.dx, kgrid.dy, getattr(kgrid, "dz", 1.0))[a] if a < nd else 1.0
    d.dt = float(kgrid.dt)
    d.Nt = int(kgrid.Nt)
    d.precision = _precision(data_cast)
    ps = _norm3(pml_size, nd, int, "pml_size", 0)
    pa = _norm3(pml_alpha, nd, _pml_alpha_cast, "pml_alpha", 0.0)
    for a in range(3):
        d.pml_size[a] = ps[a]
        d.pml_alpha[a] = pa[a]

    maps = []
    for fld, attr in (("cp", "sound_speed_compression"), ("cs", "sound_speed_shear"), ("rho0", "density")):
        m, s = _param(getattr(medium, attr), attr)
        if m is not None:
            if m.size != int(np.prod(shape)):
                raise ValueError(f"medium.{attr} must be a scalar or match the grid size {shape}")
            keep.append(m)
            setattr(d, fld, _dptr(m))
        setattr(d, fld + "_scalar", s)
        maps.append(m is not None)
    d.chi_p, d.chi_s = _kelvin_voigt(medium, not any(maps))

    if source.p0 is not None:
        p0 = _colmajor(np.broadcast_to(np.asarray(source.p0, dtype=np.float64), shape))
        keep.append(p0)
        d.p0 = _dptr(p0)

    if source.u_mask is not None:
        idx = _findall(np.broadcast_to(source.u_mask, shape))
        keep.append(idx)
        d.n_usrc = idx.size
        d.usrc_index = _iptr(idx)
        for a, nm in enumerate(("ux", "uy", "uz")[:nd]):
            sig = getattr(source, nm)
            if sig is None:
                continue
            m = _signal_rows(sig, idx.size, nm)
            flat = np.ascontiguousarray(m.ravel(order="F"))
            keep.append(flat)
            d.usrc_signal[a] = _dptr(flat)
            d.usrc_ncols[a] = m.shape[1]
        d.u_mode = int(SourceMode(source.u_mode))
    if source.s_mask is not None:
        idx = _findall(np.broadcast_to(source.s_mask, shape))
        keep.append(idx)
        d.n_ssrc = idx.size
        d.ssrc_index = _iptr(idx)
        names = ("sxx", "syy", "sxy") if nd == 2 else ("sxx", "syy", "szz", "sxy", "sxz", "syz")
        slot = {"sxx": 0, "syy": 1, "szz": 2, "sxy": 3, "sxz": 4, "syz": 5}
        for nm in names:
            sig = getattr(source, nm)
            if sig is None:
                continue
            m = _signal_rows(sig, idx.size, nm)
            flat = np.ascontiguousarray(m.ravel(order="F"))
            keep.append(flat)
            d.ssrc_signal[slot[nm]] = _dptr(flat)
            d.ssrc_ncols[slot[nm]] = m.shape[1]
        d.s_mode = int(SourceMode(source.s_mode))

    n_sensor = 0
    if sensor.mask is not None:
        m = np.asarray(sensor.mask)
        if m.dtype == bool:
            idx = _findall(m)
            keep.append(idx)
            n_sensor = idx.size
            d.sensor_index = _iptr(idx)
        else:
            n_sensor = m.shape[1]                       # Cartesian points: preallocated, never filled
        d.n_sensor = n_sensor
        bits = 0
        for f in sensor.record:
            bits |= _lib.REC_BITS[f]
        d.record = bits
    return d, keep, {"Nt": int(kgrid.Nt), "n_sensor": n_sensor}


class ElasticSimulation:
    """One GPU elastic simulation behind the C ABI (kw_elastic_create .. kw_elastic_destroy)."""

    def __init__(self, desc: "_lib.ElasticDesc", keep: List[object], device: int = 0):
        self._lib = _lib.load()
        ctx = _lib.context(device)
        h = C.c_void_p()
        _lib.check(self._lib.kw_elastic_create(ctx, C.byref(desc), C.byref(h)))
        self._h = h
        self.desc = desc
        del keep  # host inputs are borrowed only during kw_elastic_create

    @property
    def shape(self):
        return tuple(self.desc.N[a] for a in range(self.desc.ndim))

    def run(self, n_steps: int) -> None:
        _lib.check(self._lib.kw_elastic_run(self._h, int(n_steps)))

    def synchronize(self) -> None:
        _lib.check(self._lib.kw_elastic_synchronize(self._h))

    def steps_done(self) -> int:
        v = C.c_int64()
        _lib.check(self._lib.kw_elastic_steps_done(self._h, C.byref(v)))
        return v.value

    def field(self, name: str) -> np.ndarray:
        out = np.empty(int(np.prod(self.shape)), dtype=np.float64)
        _lib.check(self._lib.kw_elastic_get_field(self._h, _lib.EL_FIELD[name], out.ctypes.data_as(_DP)))
        return out.reshape(self.shape, order="F")

    def pressure(self) -> np.ndarray:
        out = np.empty(int(np.prod(self.shape)), dtype=np.float64)
        _lib.check(self._lib.kw_elastic_get_pressure(self._h, out.ctypes.data_as(_DP)))
        return out.reshape(self.shape, order="F")

    def profile(self, n_steps: int) -> Dict[str, Dict[str, float]]:
        """Run n_steps with HIP events around every launch; per-pass-class mean ms."""
        m = 32
        names = (C.c_char_p * m)()
        ms = (C.c_double * m)()
        cnt = (C.c_int64 * m)()
        nc = C.c_int32()
        _lib.check(self._lib.kw_elastic_profile(self._h, int(n_steps), m, C.byref(nc), names, ms, cnt))
        return {names[i].decode(): {"ms": ms[i], "launches": int(cnt[i])} for i in range(nc.value)}

    def sensor(self, name: str) -> np.ndarray:
        bit = _lib.REC_BITS[name]
        r, c = C.c_int64(), C.c_int64()
        _lib.check(self._lib.kw_elastic_sensor_shape(self._h, bit, C.byref(r), C.byref(c)))
        out = np.empty(r.value * c.value, dtype=np.float64)
        _lib.check(self._lib.kw_elastic_get_sensor(self._h, bit, out.ctypes.data_as(_DP)))
        if c.value == 1:
            return out
        return out.reshape((r.value, c.value), order="F")

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.check(self._lib.kw_elastic_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _run_elastic(kgrid, medium, source, sensor, pml_size, pml_alpha, data_cast, progress_callback,
                 device) -> SimulationOutput:
    desc, keep, info = build_elastic_desc(kgrid, medium, source, sensor, pml_size=pml_size, pml_alpha=pml_alpha,
                                          data_cast=data_cast)
    Nt = info["Nt"]
    with ElasticSimulation(desc, keep, device) as sim:
        if progress_callback is None:
            sim.run(Nt)
        else:
            for t in range(1, Nt + 1):
                sim.run(1)
                progress_callback(t, Nt, sim.pressure())
        sim.synchronize()
        data: Dict[str, np.ndarray] = {}
        if sensor.mask is not None:
            for f in sensor.record:
                if f in ("ux", "uy", "uz") and "xyz".index(f[1]) >= kgrid.ndim:
                    data[f] = np.zeros((info["n_sensor"], Nt))
                    continue
                data[f] = sim.sensor(f)
    return SimulationOutput(data)


def pstd_elastic_2d(kgrid: KWaveGrid2D, medium: ElasticMedium, source: ElasticSource, sensor: KWaveSensor, *,
                    pml_size=20, pml_alpha=2.0, data_cast=np.float64, plot_sim: bool = False,
                    progress_callback: Optional[Callable] = None, device: int = 0) -> SimulationOutput:
    """2-D elastic stress-velocity PSTD on the GPU (elastic.jl:133-368).  ``plot_sim``
    is accepted and unused."""
    if not isinstance(kgrid, KWaveGrid2D):
        raise TypeError("pstd_elastic_2d needs a KWaveGrid2D (MethodError)")
    return _run_elastic(kgrid, medium, source, sensor, pml_size, pml_alpha, data_cast, progress_callback, device)


def pstd_elastic_3d(kgrid: KWaveGrid3D, medium: ElasticMedium, source: ElasticSource, sensor: KWaveSensor, *,
                    pml_size=10, pml_alpha=2.0, data_cast=np.float64, plot_sim: bool = False,
                    progress_callback: Optional[Callable] = None, device: int = 0) -> SimulationOutput:
    """3-D elastic stress-velocity PSTD on the GPU (elastic.jl:386-588).  ``plot_sim``
    is accepted and unused."""
    if not isinstance(kgrid, KWaveGrid3D):
        raise TypeError("pstd_elastic_3d needs a KWaveGrid3D (MethodError)")
    return _run_elastic(kgrid, medium, source, sensor, pml_size, pml_alpha, data_cast, progress_callback, device)
