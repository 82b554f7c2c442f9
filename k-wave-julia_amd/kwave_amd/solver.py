"""kspace_first_order: drop-in for KWave.jl's solver entry point.

Mirrors ``kspace_first_order(kgrid, medium, source, sensor; kwargs...)``
(KWave.jl/src/solver/first_order.jl:92-111, 287-306, 505-524): same keyword
arguments and defaults, same validation errors, same ``SimulationOutput``
keys and Float64 shapes.  Everything after validation -- tables, the time
loop, recording and finalisation -- runs in libkwave_hip on the GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from .model import (AbstractKWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, SimulationOutput,
                    SourceMode, is_lossless, is_time_reversal, validate_record_fields)

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int64)


def _dptr(a: Optional[np.ndarray]):
    return a.ctypes.data_as(_DP) if a is not None else None


def _iptr(a: Optional[np.ndarray]):
    return a.ctypes.data_as(_IP) if a is not None else None


def _findall(mask: np.ndarray) -> np.ndarray:
    """Julia findall(mask): column-major linear indices (0-based)."""
    return np.ascontiguousarray(np.flatnonzero(np.asarray(mask, dtype=bool).ravel(order="F")), dtype=np.int64)


def _colmajor(a) -> np.ndarray:
    """Float64 copy laid out column-major (x fastest), as one flat buffer."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


def _precision(data_cast) -> int:
    if data_cast in (np.float64, float, "float64", "Float64", np.dtype("float64")):
        return 8
    if data_cast in (np.float32, "float32", "Float32", np.dtype("float32")):
        return 4
    raise TypeError(f"data_cast must be Float32 or Float64, got {data_cast!r}")


def _norm3(v, nd, cast, name, pad):
    if isinstance(v, (tuple, list)):
        if len(v) != nd:
            raise TypeError(f"{name} tuple must have {nd} entries")
        vals = [cast(x) for x in v]
    else:
        vals = [cast(v)] * nd
    return vals + [pad] * (3 - nd)


def _pml_alpha_cast(x):
    # first_order.jl:294: pml_alpha::Float64 -- an Int is a MethodError there
    if isinstance(x, (int, np.integer)) and not isinstance(x, bool):
        raise TypeError("pml_alpha must be a Float64 (an integer is not accepted)")
    return float(x)


class Simulation:
    """One GPU simulation behind the C ABI (kw_sim_create .. kw_sim_destroy).

    ``Simulation.firings(descs, keeps)`` builds the stacked form
    (kw_sim_create_firings): 2-D firings of one grid and medium that run as one
    simulation; fields then have a trailing firing axis and sensor rows are
    every firing's rows in turn."""

    def __init__(self, desc: "_lib.SimDesc", keep: List[object], device: int = 0, _firings=None):
        self._lib = _lib.load()
        ctx = _lib.context(device)
        h = C.c_void_p()
        if _firings is None:
            _lib.check(self._lib.kw_sim_create(ctx, C.byref(desc), C.byref(h)))
            self.n_firings = 1
        else:
            arr = (_lib.SimDesc * len(_firings))(*_firings)
            _lib.check(self._lib.kw_sim_create_firings(ctx, arr, len(_firings), C.byref(h)))
            self.n_firings = len(_firings)
        self._h = h
        self.desc = desc
        del keep  # host inputs are borrowed only during kw_sim_create

    @classmethod
    def firings(cls, descs: Sequence["_lib.SimDesc"], keeps: Sequence[object], device: int = 0) -> "Simulation":
        if not descs:
            raise ValueError("no firings")
        return cls(descs[0], list(keeps), device, _firings=list(descs))

    def run(self, n_steps: int) -> None:
        _lib.check(self._lib.kw_sim_run(self._h, int(n_steps)))

    def synchronize(self) -> None:
        _lib.check(self._lib.kw_sim_synchronize(self._h))

    def steps_done(self) -> int:
        v = C.c_int64()
        _lib.check(self._lib.kw_sim_steps_done(self._h, C.byref(v)))
        return v.value

    def field(self, name: str) -> np.ndarray:
        nd = self.desc.ndim
        shape = tuple(self.desc.N[a] for a in range(nd)) + ((self.n_firings,) if self.n_firings > 1 else ())
        out = np.empty(int(np.prod(shape)), dtype=np.float64)
        _lib.check(self._lib.kw_sim_get_field(self._h, _lib.FIELD[name], out.ctypes.data_as(_DP)))
        return out.reshape(shape, order="F")

    def set_field(self, name: str, value) -> None:
        """Overwrite one field of the current state (p, ux.., rhox.., rho_prev); with
        ``run(1)`` this advances one time step from a given state."""
        nd = self.desc.ndim
        shape = tuple(self.desc.N[a] for a in range(nd)) + ((self.n_firings,) if self.n_firings > 1 else ())
        a = np.asarray(value, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"field must have shape {shape}, got {a.shape}")
        buf = np.ascontiguousarray(a.ravel(order="F"))
        _lib.check(self._lib.kw_sim_set_field(self._h, _lib.FIELD[name], buf.ctypes.data_as(_DP)))

    def sensor(self, name: str) -> np.ndarray:
        bit = _lib.REC_BITS[name]
        r, c = C.c_int64(), C.c_int64()
        _lib.check(self._lib.kw_sim_sensor_shape(self._h, bit, C.byref(r), C.byref(c)))
        out = np.empty(r.value * c.value, dtype=np.float64)
        _lib.check(self._lib.kw_sim_get_sensor(self._h, bit, out.ctypes.data_as(_DP)))
        if c.value == 1:
            return out
        return out.reshape((r.value, c.value), order="F")

    def profile(self, n_steps: int) -> Dict[str, Dict[str, float]]:
        """Run n_steps with HIP events around every launch; per-kernel mean ms."""
        m = 32
        names = (C.c_char_p * m)()
        ms = (C.c_double * m)()
        cnt = (C.c_int64 * m)()
        nc = C.c_int32()
        _lib.check(self._lib.kw_sim_profile(self._h, int(n_steps), m, C.byref(nc), names, ms, cnt))
        return {names[i].decode(): {"ms": ms[i], "launches": int(cnt[i])} for i in range(nc.value)}

    def grad_c0(self, p_obs: np.ndarray, checkpoint_every: int = 50):
        """(loss, dL/dc0) of L = 1/2 sum (p - p_obs)^2 by the checkpointed adjoint."""
        shape = tuple(self.desc.N[a] for a in range(self.desc.ndim))
        obs = np.asfortranarray(np.asarray(p_obs, dtype=np.float64))
        g = np.empty(int(np.prod(shape)), dtype=np.float64)
        loss = C.c_double()
        _lib.check(self._lib.kw_sim_grad_c0(self._h, obs.ravel(order="F").ctypes.data_as(_DP), int(checkpoint_every),
                                            g.ctypes.data_as(_DP), C.byref(loss)))
        return loss.value, g.reshape(shape, order="F")

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.check(self._lib.kw_sim_destroy(self._h))
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


def quantize_map(a, storage=np.float16):
    """FP16 offset storage of a static map: returns (dequantized Float64 map, offset).

    The device stores half(a - offset) with round-to-nearest-even and
    reconstructs offset + half in the kernel precision; the dequantized map is
    what the kernels see, exactly.
    """
    a = np.asarray(a, dtype=np.float64)
    off = float(np.round((a.min() + a.max()) / 2))
    q = (a - off).astype(np.float16).astype(np.float64) + off
    return q, off


def build_desc(kgrid: AbstractKWaveGrid, medium: KWaveMedium, source: KWaveSource, sensor: KWaveSensor, *,
               pml_size=20, pml_alpha=2.0, smooth_p0: bool = True, data_cast=np.float64, medium_storage=None,
               axisymmetric: bool = False):
    """Validate like first_order.jl:311-315/529-533 and fill kw_sim_desc.

    Returns (desc, keep, info); ``keep`` holds the host buffers the
    descriptor points to (borrowed by kw_sim_create only).  ``axisymmetric``
    selects the kspace_first_order_as loop (2-D only; no time reversal).
    """
    if kgrid.Nt == 0 or kgrid.dt == 0:
        raise ValueError("Time array not set. Call make_time!(kgrid, sound_speed) before running.")
    validate_record_fields(sensor)
    nd = kgrid.ndim
    if axisymmetric and nd != 2:
        raise TypeError("kspace_first_order_as takes a 2-D grid (x axial, y radial)")
    shape = kgrid.grid_size()
    nvox = int(np.prod(shape))
    d = _lib.SimDesc()
    keep: List[object] = []
    d.ndim = nd
    for a in range(3):
        d.N[a] = shape[a] if a < nd else 1
        d.d[a] = kgrid.grid_spacing()[a] if a < nd else 1.0
    d.dt = float(kgrid.dt)
    Nt = int(kgrid.Nt)
    d.precision = _precision(data_cast)
    d.map_storage = 0
    f16 = medium_storage in (np.float16, "float16", "Float16", np.dtype("float16"))
    if medium_storage is not None and not f16:
        raise TypeError("medium_storage must be Float16 (or None for the kernel precision)")
    if f16:
        d.map_storage = 2
    ps = _norm3(pml_size, nd, int, "pml_size", 0)
    pa = _norm3(pml_alpha, nd, _pml_alpha_cast, "pml_alpha", 0.0)
    for a in range(3):
        d.pml_size[a] = ps[a]
        d.pml_alpha[a] = pa[a]

    c0 = medium.sound_speed
    rho0 = medium.density
    if np.ndim(c0) == 0:
        d.c0_scalar = float(c0)
        d.c_ref = float(c0)
    else:
        if tuple(np.shape(c0)) != tuple(shape):
            raise ValueError("sound_speed array must match the grid size")
        if f16:
            c0, d.c0_offset = quantize_map(c0)
        cm = _colmajor(c0)
        keep.append(cm)
        d.c0 = _dptr(cm)
        d.c_ref = float(np.mean(c0))                                   # first_order.jl:361
    if np.ndim(rho0) == 0:
        d.rho0_scalar = float(rho0)
    else:
        if tuple(np.shape(rho0)) != tuple(shape):
            raise ValueError("density array must match the grid size")
        if f16:
            rho0, d.rho0_offset = quantize_map(rho0)
        rm = _colmajor(rho0)
        keep.append(rm)
        d.rho0 = _dptr(rm)
    if medium.BonA is not None:
        d.nonlinear = 1
        if np.ndim(medium.BonA) == 0:
            d.bona_scalar = float(medium.BonA)
        else:
            bm = _colmajor(np.broadcast_to(medium.BonA, shape))
            keep.append(bm)
            d.bona = _dptr(bm)
    if not is_lossless(medium):
        if medium.alpha_power is None:
            raise TypeError("alpha_power must be set when alpha_coeff is set (db2neper(alpha, nothing))")
        if np.ndim(medium.alpha_coeff) != 0:
            raise TypeError("array alpha_coeff is not supported (first_order.jl:588-592 converts it to a scalar)")
        d.absorbing = 1
        d.alpha_mode = _lib.ALPHA_MODE.get(medium.alpha_mode, 2)       # any other symbol acts as :stokes
        d.alpha_coeff = float(medium.alpha_coeff)
        d.alpha_power = float(medium.alpha_power)

    d.axisymmetric = 1 if axisymmetric else 0
    tr = is_time_reversal(sensor) and not axisymmetric      # axisymmetric.jl has no time reversal
    if tr:
        trd = np.asarray(sensor.time_reversal_boundary_data, dtype=np.float64)
        Nt = int(trd.shape[1])                                          # first_order.jl:430-431
        kgrid.Nt = Nt
    d.Nt = Nt

    if source.p0 is not None and not tr:
        if tuple(source.p0.shape) != tuple(shape):
            raise ValueError("p0 must match the grid size")
        p0 = _colmajor(source.p0)
        keep.append(p0)
        d.p0 = _dptr(p0)
        if smooth_p0:
            ok = (shape[0] > 4) if nd == 1 else all(n > 1 for n in shape)   # first_order_steps.jl:70,109,157
            d.smooth_p0 = 1 if ok else 0

    if source.p_mask is not None and source.p is not None:
        idx = _findall(source.p_mask)
        sig = np.asarray(source.p, dtype=np.float64)
        if sig.ndim == 1:
            sig = sig[:, None]                                          # Vector: size(p, 2) == 1
        if sig.shape[0] < idx.size:
            raise IndexError("source.p has fewer rows than p_mask has points")
        sig = np.asfortranarray(sig[: idx.size])
        keep += [idx, sig]
        d.n_psrc = idx.size
        d.psrc_index = _iptr(idx)
        d.psrc_signal = sig.ctypes.data_as(_DP)
        d.psrc_ncols = sig.shape[1]
        d.p_mode = int(source.p_mode)
    if source.u_mask is not None:
        idx = _findall(source.u_mask)
        keep.append(idx)
        d.n_usrc = idx.size
        d.usrc_index = _iptr(idx)
        for a, s in enumerate((source.ux, source.uy, source.uz)[:nd]):
            if s is None:
                continue
            s = np.asarray(s, dtype=np.float64)
            if s.ndim == 1:
                s = s[:, None]
            if s.shape[0] < idx.size:
                raise IndexError("velocity signal has fewer rows than u_mask has points")
            s = np.asfortranarray(s[: idx.size])
            keep.append(s)
            d.usrc_signal[a] = s.ctypes.data_as(_DP)
            d.usrc_ncols[a] = s.shape[1]
        d.u_mode = int(source.u_mode)

    info = {"n_sensor": 0, "cartesian": False, "tr": tr, "Nt": Nt}
    rec = 0
    for f in sensor.record:
        rec |= _lib.REC_BITS[f]
    d.record = rec
    if sensor.mask is not None:
        m = sensor.mask
        if m.dtype == bool:
            idx = _findall(m)
            keep.append(idx)
            d.n_sensor = idx.size
            d.sensor_index = _iptr(idx) if idx.size else None
            info["n_sensor"] = idx.size
        else:
            d.n_sensor = m.shape[1]                                     # Cartesian: never filled (A-13)
            d.sensor_index = None
            info["n_sensor"] = m.shape[1]
            info["cartesian"] = True
    if tr:
        # first_order.jl:455 reads tr_data[j, tr_t] for every sensor point j: too few rows is a BoundsError
        if trd.ndim != 2 or trd.shape[0] < info["n_sensor"]:
            raise IndexError("time_reversal_boundary_data has fewer rows than the sensor mask has points")
        trc = np.asfortranarray(trd)
        keep.append(trc)
        d.tr_data = trc.ctypes.data_as(_DP)
    if sensor.directivity_angle is not None and sensor.directivity_size is not None:
        d.directivity = 1
        ang = np.asarray(sensor.directivity_angle, dtype=np.float64)
        if ang.ndim == 0:
            d.directivity_angle_scalar = float(ang)
        else:
            ang = np.ascontiguousarray(ang.ravel(order="F"))
            if ang.size < info["n_sensor"]:
                raise IndexError("directivity_angle has fewer entries than sensor points")
            keep.append(ang)
            d.directivity_angle = _dptr(ang)
        d.directivity_size = float(sensor.directivity_size)
    if sensor.frequency_response is not None:
        d.freq_response = 1
        d.freq_center, d.freq_bandwidth = sensor.frequency_response
    return d, keep, info


def _save_to_disk(path: str, kgrid: AbstractKWaveGrid, medium: KWaveMedium, source: KWaveSource,
                  sensor: KWaveSensor, pml_size, pml_alpha) -> SimulationOutput:
    """Save-to-disk mode (first_order.jl:129-140 1-D, 329-345 2-D, 546-562 3-D): after the
    input validation, write the k-Wave C++ input file (attributes, grid, c0 or c_ref,
    rho0 for 2-D/3-D maps, p0) and return an empty SimulationOutput without running."""
    from . import hdf5
    if kgrid.Nt == 0 or kgrid.dt == 0:
        raise ValueError("Time array not set. Call make_time!(kgrid, sound_speed) before running.")
    validate_record_fields(sensor)
    nd = kgrid.ndim
    sizes = tuple(pml_size) if isinstance(pml_size, (tuple, list)) else (pml_size,) * nd
    alphas = tuple(pml_alpha) if isinstance(pml_alpha, (tuple, list)) else (pml_alpha,) * nd
    hdf5.write_attributes(path)
    hdf5.write_grid(path, kgrid, pml_size=sizes, pml_alpha=alphas)
    if np.ndim(medium.sound_speed) > 0:
        hdf5.write_matrix(path, "c0", np.asarray(medium.sound_speed, dtype=np.float64))
    else:
        hdf5.write_matrix(path, "c_ref", float(medium.sound_speed))
    if nd > 1 and np.ndim(medium.density) > 0:
        hdf5.write_matrix(path, "rho0", np.asarray(medium.density, dtype=np.float64))
    if source.p0 is not None:
        hdf5.write_matrix(path, "p0", np.asarray(source.p0, dtype=np.float64))
    return SimulationOutput({})


def kspace_first_order(kgrid: AbstractKWaveGrid, medium: KWaveMedium, source: KWaveSource,
                       sensor: KWaveSensor, *, pml_inside: bool = True, pml_size=20, pml_alpha=2.0,
                       smooth_p0: bool = True, smooth_c0: bool = False, smooth_rho0: bool = False,
                       data_cast=np.float64, save_to_disk: Optional[str] = None, plot_sim: bool = False,
                       plot_layout: str = "default", plot_scale="auto", record_movie: Optional[str] = None,
                       progress_callback: Optional[Callable] = None, show_progress: bool = True,
                       medium_storage=None, device: int = 0) -> SimulationOutput:
    """Run a k-space PSTD simulation on the GPU (first_order.jl:92/287/505).

    ``pml_inside``, ``smooth_c0``, ``smooth_rho0``, ``plot_*`` and
    ``record_movie`` are accepted and unused, as in the reference.
    ``medium_storage=np.float16`` (not in the reference) keeps the c0 and
    rho0 maps in HBM as FP16 offsets; arithmetic stays in ``data_cast``.
    """
    if save_to_disk is not None:
        return _save_to_disk(save_to_disk, kgrid, medium, source, sensor, pml_size, pml_alpha)
    desc, keep, info = build_desc(kgrid, medium, source, sensor, pml_size=pml_size, pml_alpha=pml_alpha,
                                  smooth_p0=smooth_p0, data_cast=data_cast, medium_storage=medium_storage)
    return _run(desc, keep, info, kgrid, sensor, progress_callback, device)


def kspace_first_order_as(kgrid: AbstractKWaveGrid, medium: KWaveMedium, source: KWaveSource,
                          sensor: KWaveSensor, *, pml_inside: bool = True, pml_size=20, pml_alpha=2.0,
                          smooth_p0: bool = True, smooth_c0: bool = False, smooth_rho0: bool = False,
                          data_cast=np.float64, plot_sim: bool = False, plot_layout: str = "default",
                          plot_scale="auto", record_movie: Optional[str] = None,
                          progress_callback: Optional[Callable] = None, medium_storage=None,
                          device: int = 0) -> SimulationOutput:
    """Axisymmetric k-space PSTD on the GPU (axisymmetric.jl:26-227).

    A 2-D grid with x axial and y radial (r = |y|); the radial density update
    carries the ur/r term of the cylindrical divergence.  Keywords as in the
    reference (``pml_inside``, ``smooth_c0``, ``smooth_rho0``, ``plot_*``,
    ``record_movie`` accepted and unused).
    """
    desc, keep, info = build_desc(kgrid, medium, source, sensor, pml_size=pml_size, pml_alpha=pml_alpha,
                                  smooth_p0=smooth_p0, data_cast=data_cast, medium_storage=medium_storage,
                                  axisymmetric=True)
    return _run(desc, keep, info, kgrid, sensor, progress_callback, device)


def kspace_first_order_firings(kgrid: AbstractKWaveGrid, medium: KWaveMedium, sources: Sequence[KWaveSource],
                               sensors, *, pml_size=20, pml_alpha=2.0, data_cast=np.float64, medium_storage=None,
                               device: int = 0, **ignored) -> List[SimulationOutput]:
    """Transmit sequence: one 2-D simulation per source (firing) on one grid and medium,
    run as one stacked simulation on the GPU (kw_sim_create_firings).

    Returns one SimulationOutput per firing, equal to
    ``kspace_first_order(kgrid, medium, sources[i], sensors[i], ...)``.  ``sensors`` is one
    sensor for all firings or one per firing (binary masks).  Time-varying sources only:
    p0, time reversal and sensor post-filters are not part of this form.
    """
    if kgrid.ndim != 2:
        raise TypeError("kspace_first_order_firings takes a 2-D grid")
    sources = list(sources)
    sens = list(sensors) if isinstance(sensors, (list, tuple)) else [sensors] * len(sources)
    if len(sens) != len(sources) or not sources:
        raise ValueError("one sensor per firing (or one sensor for all) and at least one firing")
    descs, keeps, infos = [], [], []
    for src, sen in zip(sources, sens):
        if src.p0 is not None:
            raise ValueError("kspace_first_order_firings takes time-varying sources (no p0)")
        d, k, info = build_desc(kgrid, medium, src, sen, pml_size=pml_size, pml_alpha=pml_alpha, smooth_p0=False,
                                data_cast=data_cast, medium_storage=medium_storage)
        if info["cartesian"] or info["tr"]:
            raise ValueError("kspace_first_order_firings needs binary sensor masks and no time reversal")
        descs.append(d)
        keeps.append(k)
        infos.append(info)
    Nt = infos[0]["Nt"]
    with Simulation.firings(descs, keeps, device) as sim:
        sim.run(Nt)
        sim.synchronize()
        outs = [dict() for _ in sources]
        record = list(sens[0].record) if sens[0].mask is not None else []
        for f in record:
            if f in ("uz",):
                for i, info in enumerate(infos):
                    outs[i][f] = np.zeros((info["n_sensor"], Nt))
                continue
            whole = sim.sensor(f)
            row = 0
            for i, info in enumerate(infos):
                n = info["n_sensor"]
                outs[i][f] = whole[row:row + n]
                row += n
    return [SimulationOutput(o) for o in outs]


def _run(desc, keep, info, kgrid, sensor, progress_callback, device) -> SimulationOutput:
    Nt = info["Nt"]
    with Simulation(desc, keep, device) as sim:
        if progress_callback is None:
            sim.run(Nt)
        else:
            for t in range(1, Nt + 1):
                sim.run(1)
                progress_callback(t, Nt, sim.field("p"))
        sim.synchronize()
        data: Dict[str, np.ndarray] = {}
        if sensor.mask is not None:
            for f in sensor.record:
                if f in ("ux", "uy", "uz") and "xyz".index(f[1]) >= kgrid.ndim:
                    data[f] = np.zeros((info["n_sensor"], Nt))
                    continue
                data[f] = sim.sensor(f)
        if info["tr"]:
            data["p_final"] = sim.sensor("p_final") if "p_final" in sensor.record else sim.field("p").ravel(order="F")
    return SimulationOutput(data)
