"""Linear-array transducer (transducer.jl): element masks on a 3-D grid,
focused/steered transmit signals, receive combination with apodization.

Host-side preparation and post-processing around the simulation: the mask
and signal matrix become ``KWaveSource`` fields; ``combine_transducer_sensor_data``
reduces recorded rows per element.  Positions, widths and lengths are grid
points (1-based), elements are laid out along y with a kerf
``element_spacing``, and span ``element_length`` points along z from
``position``.

Row order of the transmit signal matrix follows the reference: element by
element, z outer and y inner within an element.  The solver pairs row j with
the j-th point of ``findall(mask)`` (x, then y, then z fastest-first), so for
several multi-row elements the rows land on points in that order, as they do
in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable, List, Optional, Tuple

import numpy as np

from .model import KWaveGrid3D


@dataclass
class KWaveTransducer:
    """transducer.jl:13-26 (keyword construction, same defaults)."""
    number_elements: int
    element_width: int
    element_length: int
    element_spacing: int = 0
    position: Tuple[int, int, int] = (1, 1, 1)
    radius: float = float("inf")
    focus_distance: float = float("inf")
    steering_angle: float = 0.0
    transmit_apodization: str = "rectangular"
    receive_apodization: str = "rectangular"
    active_elements: Optional[List[int]] = None
    input_signal: Optional[np.ndarray] = None


def _active(t: KWaveTransducer) -> List[int]:
    return list(range(1, t.number_elements + 1)) if t.active_elements is None else [int(e) for e in t.active_elements]


def _element_points(t: KWaveTransducer, kgrid, elem: int):
    """Grid points (1-based x, y, z) of one element in the reference's loop order
    (z outer, y inner), clipped to the grid (transducer.jl:54-68)."""
    px, py, pz = t.position
    y0 = py + (elem - 1) * (t.element_width + t.element_spacing)
    y1 = y0 + t.element_width - 1
    z0, z1 = pz, pz + t.element_length - 1
    if not 1 <= px <= kgrid.Nx:
        return []
    return [(px, j, k) for k in range(max(1, z0), min(kgrid.Nz, z1) + 1)
            for j in range(max(1, y0), min(kgrid.Ny, y1) + 1)]


def _check3d(kgrid):
    if not isinstance(kgrid, KWaveGrid3D):
        raise TypeError("the transducer functions need a KWaveGrid3D (MethodError)")


def get_transducer_binary_mask(transducer: KWaveTransducer, kgrid) -> np.ndarray:
    """transducer.jl:44-71."""
    _check3d(kgrid)
    mask = np.zeros((kgrid.Nx, kgrid.Ny, kgrid.Nz), dtype=bool)
    for e in _active(transducer):
        for (i, j, k) in _element_points(transducer, kgrid, e):
            mask[i - 1, j - 1, k - 1] = True
    return mask


def _window(n: int, kind: str) -> np.ndarray:
    """Symmetric windows of get_win (filters.jl:183-220) used for apodization."""
    if n <= 0:
        return np.zeros(0)
    if n == 1:
        return np.ones(1)
    x = 2.0 * np.pi * np.arange(n) / (n - 1)
    if kind in ("hann", "hanning"):
        return 0.5 * (1.0 - np.cos(x))
    if kind == "hamming":
        return 0.54 - 0.46 * np.cos(x)
    if kind == "blackman":
        return 0.42 - 0.5 * np.cos(x) + 0.08 * np.cos(2.0 * x)
    if kind in ("triangular", "bartlett"):
        return 1.0 - np.abs(2.0 * np.arange(n) / (n - 1) - 1.0)
    raise ValueError(f"Unknown apodization type: :{kind}")


def compute_apodization(n_elements: int, kind: str) -> np.ndarray:
    """transducer.jl:231-245."""
    kind = str(kind).lstrip(":")
    if kind == "rectangular":
        return np.ones(n_elements)
    return _window(n_elements, kind)


def transmit_delays(transducer: KWaveTransducer, kgrid, c0: float) -> np.ndarray:
    """Per-active-element delays [s] (transducer.jl:99-141): focusing
    (sqrt(F^2 + dy^2) - F)/c0 about the mean element centre, plus steering
    dy sin(theta)/c0, shifted so the smallest is zero."""
    act = _active(transducer)
    px, py, pz = transducer.position
    yc = np.array([py + (e - 1) * (transducer.element_width + transducer.element_spacing)
                   + transducer.element_width / 2.0 for e in act], dtype=np.float64)
    d = np.zeros(len(act))
    if len(act) == 0:
        return d
    dy = (yc - yc.mean()) * kgrid.dy
    if np.isfinite(transducer.focus_distance):
        F = float(transducer.focus_distance)
        d += (np.sqrt(F * F + dy * dy) - F) / c0
    if transducer.steering_angle != 0.0:
        d += dy * np.sin(transducer.steering_angle) / c0
    return d - d.min()


def get_transducer_source(transducer: KWaveTransducer, kgrid, medium) -> Tuple[np.ndarray, np.ndarray]:
    """transducer.jl:87-170 -- (mask, n_points x Nt signals): the input signal
    delayed by round(delay/dt) samples and scaled by the transmit apodization."""
    _check3d(kgrid)
    if transducer.input_signal is None:
        raise ValueError("Transducer input_signal must be set before generating source")
    mask = get_transducer_binary_mask(transducer, kgrid)
    n_points = int(mask.sum())
    Nt, dt = int(kgrid.Nt), float(kgrid.dt)
    c = medium.sound_speed
    c0 = float(c) if np.ndim(c) == 0 else float(np.mean(c))
    act = _active(transducer)
    apod = compute_apodization(len(act), transducer.transmit_apodization)
    delays = transmit_delays(transducer, kgrid, c0)
    sig = np.asarray(transducer.input_signal, dtype=np.float64).ravel()
    out = np.zeros((n_points, Nt))
    row = 0
    for ei, e in enumerate(act):
        shift = int(round(delays[ei] / dt))
        n = max(0, min(sig.size, Nt - shift))
        for _ in _element_points(transducer, kgrid, e):
            if row < n_points:               # overlapping elements: the mask holds each point once
                out[row, shift:shift + n] = sig[:n] * apod[ei]
            row += 1
    if row > n_points:
        raise IndexError("overlapping transducer elements: more element points than mask points (BoundsError)")
    return mask, out


def combine_transducer_sensor_data(transducer: KWaveTransducer, kgrid, sensor_data) -> np.ndarray:
    """transducer.jl:186-229 -- per-active-element mean of consecutive rows (element
    loop order) times the receive apodization."""
    _check3d(kgrid)
    data = np.asarray(sensor_data, dtype=np.float64)
    act = _active(transducer)
    apod = compute_apodization(len(act), transducer.receive_apodization)
    out = np.zeros((len(act), data.shape[1]))
    row = 0
    for ei, e in enumerate(act):
        npts = 0
        for _ in _element_points(transducer, kgrid, e):
            if row < data.shape[0]:
                out[ei] += data[row]
                npts += 1
            row += 1
        if npts:
            out[ei] *= apod[ei] / npts
    return out
