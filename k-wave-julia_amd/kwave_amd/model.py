"""Data model mirroring KWave.jl's grid, medium, source, sensor and output.

Names, fields and defaults follow the reference so scripts translate one to
one (Julia ``make_time!(kgrid, c)`` -> Python ``make_time(kgrid, c)``,
``kgrid.dt[]`` -> ``kgrid.dt``, symbols ``:p`` -> strings ``"p"``).  Arrays
are NumPy arrays indexed ``[i, j, k]`` like the reference's ``[i+1, j+1,
k+1]``; the reference's column-major order is NumPy's Fortran order.
"""
from __future__ import annotations

import enum
import math
from typing import Dict, Iterable, Optional, Sequence, Tuple, Union

import numpy as np


class SourceMode(enum.IntEnum):
    """types.jl:14 -- same order as the reference enum."""
    Dirichlet = 0
    Additive = 1
    AdditiveNoCorrection = 2


Dirichlet = SourceMode.Dirichlet
Additive = SourceMode.Additive
AdditiveNoCorrection = SourceMode.AdditiveNoCorrection


# ---------------------------------------------------------------------------
# Grid (grid.jl)
# ---------------------------------------------------------------------------

def _make_wavenumber_vec(N: int, d: float) -> np.ndarray:
    """grid.jl:120-130."""
    dk = 2 * math.pi / (N * d)
    if N % 2 == 0:
        m = np.concatenate([np.arange(0, N // 2), np.arange(-(N // 2), 0)])
    else:
        m = np.concatenate([np.arange(0, (N - 1) // 2 + 1), np.arange(-((N - 1) // 2), 0)])
    return m.astype(np.float64) * dk


def _make_spatial_vec(N: int, d: float) -> np.ndarray:
    """grid.jl:137-139."""
    return (np.arange(N) - (N - 1) / 2.0) * d


def _make_shift_operators(kv: np.ndarray, d: float):
    """grid.jl:149-153."""
    return np.exp(1j * kv * d / 2), np.exp(-1j * kv * d / 2)


class AbstractKWaveGrid:
    """Common behaviour of KWaveGrid1D/2D/3D (grid.jl:251-296)."""
    ndim: int

    def __init__(self):
        self.dt = 0.0
        self.Nt = 0
        self.t_array = np.zeros(0)

    def grid_size(self) -> Tuple[int, ...]:
        return tuple(getattr(self, n) for n in ("Nx", "Ny", "Nz")[: self.ndim])

    def grid_spacing(self) -> Tuple[float, ...]:
        return tuple(getattr(self, n) for n in ("dx", "dy", "dz")[: self.ndim])

    def total_grid_points(self) -> int:
        return int(np.prod(self.grid_size()))

    def k_max(self) -> float:
        return min(math.pi / d for d in self.grid_spacing())


class KWaveGrid1D(AbstractKWaveGrid):
    ndim = 1

    def __init__(self, Nx: int, dx: float):
        super().__init__()
        self.Nx, self.dx = int(Nx), float(dx)
        self.x_vec = _make_spatial_vec(self.Nx, self.dx)
        self.kx_vec = _make_wavenumber_vec(self.Nx, self.dx)
        self.k = np.abs(self.kx_vec)
        self.ddx_k_shift_pos, self.ddx_k_shift_neg = _make_shift_operators(self.kx_vec, self.dx)


class KWaveGrid2D(AbstractKWaveGrid):
    ndim = 2

    def __init__(self, Nx: int, dx: float, Ny: int, dy: float):
        super().__init__()
        self.Nx, self.dx, self.Ny, self.dy = int(Nx), float(dx), int(Ny), float(dy)
        self.x_vec = _make_spatial_vec(self.Nx, self.dx)
        self.y_vec = _make_spatial_vec(self.Ny, self.dy)
        self.kx_vec = _make_wavenumber_vec(self.Nx, self.dx)
        self.ky_vec = _make_wavenumber_vec(self.Ny, self.dy)
        self.k = np.sqrt(self.kx_vec[:, None] ** 2 + self.ky_vec[None, :] ** 2)
        self.ddx_k_shift_pos, self.ddx_k_shift_neg = _make_shift_operators(self.kx_vec, self.dx)
        self.ddy_k_shift_pos, self.ddy_k_shift_neg = _make_shift_operators(self.ky_vec, self.dy)


class KWaveGrid3D(AbstractKWaveGrid):
    ndim = 3

    def __init__(self, Nx: int, dx: float, Ny: int, dy: float, Nz: int, dz: float):
        super().__init__()
        self.Nx, self.dx, self.Ny, self.dy = int(Nx), float(dx), int(Ny), float(dy)
        self.Nz, self.dz = int(Nz), float(dz)
        self.x_vec = _make_spatial_vec(self.Nx, self.dx)
        self.y_vec = _make_spatial_vec(self.Ny, self.dy)
        self.z_vec = _make_spatial_vec(self.Nz, self.dz)
        self.kx_vec = _make_wavenumber_vec(self.Nx, self.dx)
        self.ky_vec = _make_wavenumber_vec(self.Ny, self.dy)
        self.kz_vec = _make_wavenumber_vec(self.Nz, self.dz)
        self.ddx_k_shift_pos, self.ddx_k_shift_neg = _make_shift_operators(self.kx_vec, self.dx)
        self.ddy_k_shift_pos, self.ddy_k_shift_neg = _make_shift_operators(self.ky_vec, self.dy)
        self.ddz_k_shift_pos, self.ddz_k_shift_neg = _make_shift_operators(self.kz_vec, self.dz)

    @property
    def k(self) -> np.ndarray:
        # built on demand: a 512^3 Float64 |k| is 1 GiB and the solver never needs it
        return np.sqrt(self.kx_vec[:, None, None] ** 2 + self.ky_vec[None, :, None] ** 2
                       + self.kz_vec[None, None, :] ** 2)


def KWaveGrid(*args):
    """KWaveGrid(Nx, dx[, Ny, dy[, Nz, dz]]) (grid.jl:183-245)."""
    if len(args) == 2:
        return KWaveGrid1D(*args)
    if len(args) == 4:
        return KWaveGrid2D(*args)
    if len(args) == 6:
        return KWaveGrid3D(*args)
    raise TypeError("KWaveGrid takes (Nx, dx), (Nx, dx, Ny, dy) or (Nx, dx, Ny, dy, Nz, dz)")


def make_time(grid: AbstractKWaveGrid, sound_speed, cfl: float = 0.3, t_end: Optional[float] = None):
    """make_time!(grid, c; cfl=0.3, t_end=nothing) (grid.jl:320-352)."""
    c = np.asarray(sound_speed, dtype=np.float64)
    c_max, c_min = float(np.max(c)), float(np.min(c))
    ds = grid.grid_spacing()
    dt = cfl * min(ds) / c_max
    grid.dt = dt
    if t_end is None:
        diag = math.sqrt(sum((n * d) ** 2 for n, d in zip(grid.grid_size(), ds)))
        t_end = 2.0 * diag / c_min
    Nt = int(math.ceil(t_end / dt)) + 1
    grid.Nt = Nt
    grid.t_array = np.arange(Nt) * dt
    return dt, Nt


# ---------------------------------------------------------------------------
# Medium (medium.jl)
# ---------------------------------------------------------------------------

def _f64(x):
    if x is None:
        return None
    a = np.asarray(x, dtype=np.float64)
    return float(a) if a.ndim == 0 else a


class KWaveMedium:
    """KWaveMedium(; sound_speed, density=1.0, alpha_coeff, alpha_power,
    alpha_mode=:no_absorption, BonA) (medium.jl:18-76); inputs become Float64."""

    def __init__(self, *, sound_speed, density=1.0, alpha_coeff=None, alpha_power=None,
                 alpha_mode: str = "no_absorption", BonA=None):
        self.sound_speed = _f64(sound_speed)
        self.density = _f64(density)
        self.alpha_coeff = _f64(alpha_coeff)
        self.alpha_power = None if alpha_power is None else float(alpha_power)
        self.alpha_mode = str(alpha_mode).lstrip(":")
        self.BonA = _f64(BonA)


def is_homogeneous(m: KWaveMedium) -> bool:
    return np.ndim(m.sound_speed) == 0 and np.ndim(m.density) == 0


def is_lossless(m: KWaveMedium) -> bool:
    """medium.jl:90."""
    return m.alpha_coeff is None or m.alpha_mode == "no_absorption"


def is_nonlinear(m: KWaveMedium) -> bool:
    return m.BonA is not None


# ---------------------------------------------------------------------------
# Source (source.jl)
# ---------------------------------------------------------------------------

class KWaveSource:
    """KWaveSource(; p0, p_mask, p, p_mode=Additive, u_mask, ux, uy, uz,
    u_mode=Additive) (source.jl:24-73)."""

    def __init__(self, *, p0=None, p_mask=None, p=None, p_mode: SourceMode = Additive, u_mask=None,
                 ux=None, uy=None, uz=None, u_mode: SourceMode = Additive):
        f = lambda x: None if x is None else np.asarray(x, dtype=np.float64)
        b = lambda x: None if x is None else np.asarray(x, dtype=bool)
        self.p0 = f(p0)
        self.p_mask = b(p_mask)
        self.p = f(p)
        self.p_mode = SourceMode(p_mode)
        self.u_mask = b(u_mask)
        self.ux, self.uy, self.uz = f(ux), f(uy), f(uz)
        self.u_mode = SourceMode(u_mode)


class ElasticMedium:
    """ElasticMedium(; sound_speed_compression, sound_speed_shear, density=1000,
    alpha_coeff_compression, alpha_coeff_shear, alpha_power) (elastic.jl:20-48)."""

    def __init__(self, *, sound_speed_compression, sound_speed_shear, density=1000.0,
                 alpha_coeff_compression=None, alpha_coeff_shear=None, alpha_power=None):
        self.sound_speed_compression = _f64(sound_speed_compression)
        self.sound_speed_shear = _f64(sound_speed_shear)
        self.density = _f64(density)
        self.alpha_coeff_compression = _f64(alpha_coeff_compression)
        self.alpha_coeff_shear = _f64(alpha_coeff_shear)
        self.alpha_power = None if alpha_power is None else float(alpha_power)


class ElasticSource:
    """ElasticSource(; s_mask, sxx, syy, szz, sxy, sxz, syz, s_mode=Additive, u_mask,
    ux, uy, uz, u_mode=Additive, p0) (elastic.jl:57-103)."""

    def __init__(self, *, s_mask=None, sxx=None, syy=None, szz=None, sxy=None, sxz=None, syz=None,
                 s_mode: SourceMode = Additive, u_mask=None, ux=None, uy=None, uz=None,
                 u_mode: SourceMode = Additive, p0=None):
        f = lambda x: None if x is None else np.asarray(x, dtype=np.float64)
        b = lambda x: None if x is None else np.asarray(x, dtype=bool)
        self.s_mask = b(s_mask)
        self.sxx, self.syy, self.szz = f(sxx), f(syy), f(szz)
        self.sxy, self.sxz, self.syz = f(sxy), f(sxz), f(syz)
        self.s_mode = SourceMode(s_mode)
        self.u_mask = b(u_mask)
        self.ux, self.uy, self.uz = f(ux), f(uy), f(uz)
        self.u_mode = SourceMode(u_mode)
        self.p0 = f(p0)


def has_p0(s: KWaveSource) -> bool:
    return s.p0 is not None


def has_pressure_source(s: KWaveSource) -> bool:
    return s.p_mask is not None and s.p is not None


def has_velocity_source(s: KWaveSource) -> bool:
    return s.u_mask is not None


# ---------------------------------------------------------------------------
# Sensor and output (sensor.jl)
# ---------------------------------------------------------------------------

VALID_RECORD_FIELDS = frozenset([                                      # sensor.jl:41-45
    "p", "p_max", "p_min", "p_rms", "p_final",
    "ux", "uy", "uz", "u_max", "u_rms", "u_final",
    "I_avg", "I_max",
])


class KWaveSensor:
    """KWaveSensor(; mask=nothing, record=[:p], time_reversal_boundary_data,
    directivity_angle, directivity_size, frequency_response) (sensor.jl:31-38)."""

    def __init__(self, *, mask=None, record: Iterable[str] = ("p",), time_reversal_boundary_data=None,
                 directivity_angle=None, directivity_size=None, frequency_response=None):
        if mask is not None:
            mask = np.asarray(mask)
            if mask.dtype != bool:
                mask = mask.astype(np.float64)
        self.mask = mask
        self.record = [str(r).lstrip(":") for r in record]
        self.time_reversal_boundary_data = (None if time_reversal_boundary_data is None
                                            else np.asarray(time_reversal_boundary_data, dtype=np.float64))
        self.directivity_angle = directivity_angle
        self.directivity_size = None if directivity_size is None else float(directivity_size)
        self.frequency_response = (None if frequency_response is None
                                   else (float(frequency_response[0]), float(frequency_response[1])))


def validate_record_fields(sensor: KWaveSensor) -> None:
    """sensor.jl:52-58."""
    for f in sensor.record:
        if f not in VALID_RECORD_FIELDS:
            raise ValueError(f"Invalid sensor record field: :{f}. Valid fields: "
                             + ", ".join(sorted(VALID_RECORD_FIELDS)))


def is_time_reversal(s: KWaveSensor) -> bool:
    return s.time_reversal_boundary_data is not None


class SimulationOutput:
    """sensor.jl:76-87 -- dict-like container of Float64 results."""

    def __init__(self, data: Dict[str, np.ndarray]):
        self.data = data

    def __getitem__(self, key: str) -> np.ndarray:
        return self.data[str(key).lstrip(":")]

    def __contains__(self, key: str) -> bool:
        return str(key).lstrip(":") in self.data

    def haskey(self, key: str) -> bool:
        return key in self

    def keys(self):
        return self.data.keys()

    def __repr__(self) -> str:
        return "SimulationOutput with fields: " + ", ".join(sorted(self.data))
