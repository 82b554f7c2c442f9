"""Off-grid transducer arrays (array.jl): elements in physical coordinates,
mapped to grid masks, with per-element source signals spread over each
element's grid points and recorded sensor data averaged back per element.

Host-side preparation and post-processing around the simulation: the
combined mask and the distributed signal matrix become ``KWaveSource``
fields, and ``combine_sensor_data`` reduces the recorded ``p`` rows.
Element indices are 1-based, as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple, Union

import numpy as np

from .builders import make_arc, make_ball, make_bowl
from .model import KWaveGrid2D, KWaveGrid3D


@dataclass(frozen=True)
class ArcElement:        # array.jl:19-24
    position: Tuple[float, float]
    radius: float
    diameter: float
    focus: Tuple[float, float]


@dataclass(frozen=True)
class BowlElement:       # array.jl:27-32
    position: Tuple[float, float, float]
    radius: float
    diameter: float
    focus: Tuple[float, float, float]


@dataclass(frozen=True)
class DiscElement:       # array.jl:35-39
    position: Tuple[float, float, float]
    diameter: float
    focus: Tuple[float, float, float]


@dataclass(frozen=True)
class RectElement:       # array.jl:42-47
    position: Tuple[float, float, float]
    width: float
    height: float
    focus: Tuple[float, float, float]


Element = Union[ArcElement, BowlElement, DiscElement, RectElement]


class KWaveArray:
    """array.jl:68-72 -- an ordered list of elements."""

    def __init__(self, elements: List[Element] | None = None):
        self.elements: List[Element] = list(elements or [])

    def __len__(self) -> int:
        return len(self.elements)


def _f(v) -> Tuple[float, ...]:
    return tuple(float(x) for x in v)


def add_arc_element(array: KWaveArray, position, radius, diameter, focus) -> KWaveArray:
    """array.jl:84-89."""
    array.elements.append(ArcElement(_f(position), float(radius), float(diameter), _f(focus)))
    return array


def add_bowl_element(array: KWaveArray, position, radius, diameter, focus) -> KWaveArray:
    """array.jl:96-101."""
    array.elements.append(BowlElement(_f(position), float(radius), float(diameter), _f(focus)))
    return array


def add_disc_element(array: KWaveArray, position, diameter, focus) -> KWaveArray:
    """array.jl:108-113."""
    array.elements.append(DiscElement(_f(position), float(diameter), _f(focus)))
    return array


def add_rect_element(array: KWaveArray, position, width, height, focus) -> KWaveArray:
    """array.jl:120-125."""
    array.elements.append(RectElement(_f(position), float(width), float(height), _f(focus)))
    return array


def _nearest(vec: np.ndarray, v: float) -> int:
    """argmin(abs.(vec .- v)) -- first minimum, 1-based."""
    return int(np.argmin(np.abs(vec - v))) + 1


def _element_mask(elem: Element, kgrid) -> np.ndarray:
    """array.jl:157-208."""
    if isinstance(elem, ArcElement):
        if not isinstance(kgrid, KWaveGrid2D):
            raise TypeError("arc elements need a 2-D grid")
        cx, cy = _nearest(kgrid.x_vec, elem.position[0]), _nearest(kgrid.y_vec, elem.position[1])
        fx, fy = _nearest(kgrid.x_vec, elem.focus[0]), _nearest(kgrid.y_vec, elem.focus[1])
        return make_arc(kgrid.Nx, kgrid.Ny, cx, cy, elem.radius / kgrid.dx, elem.diameter / kgrid.dx, (fx, fy))
    if not isinstance(kgrid, KWaveGrid3D):
        raise TypeError(f"{type(elem).__name__} needs a 3-D grid")
    c = (_nearest(kgrid.x_vec, elem.position[0]), _nearest(kgrid.y_vec, elem.position[1]),
         _nearest(kgrid.z_vec, elem.position[2]))
    if isinstance(elem, BowlElement):
        f = (_nearest(kgrid.x_vec, elem.focus[0]), _nearest(kgrid.y_vec, elem.focus[1]),
             _nearest(kgrid.z_vec, elem.focus[2]))
        return make_bowl(kgrid.Nx, kgrid.Ny, kgrid.Nz, c, elem.radius / kgrid.dx, elem.diameter / kgrid.dx, f)
    if isinstance(elem, DiscElement):
        # the reference maps a disc element to a solid ball of the disc's radius
        return make_ball(kgrid.Nx, kgrid.Ny, kgrid.Nz, c[0], c[1], c[2], elem.diameter / (2 * kgrid.dx))
    if isinstance(elem, RectElement):
        hw = int(np.rint(elem.width / (2 * kgrid.dx)))
        hh = int(np.rint(elem.height / (2 * kgrid.dy)))
        mask = np.zeros((kgrid.Nx, kgrid.Ny, kgrid.Nz), dtype=bool)
        if 1 <= c[2] <= kgrid.Nz:
            i0, i1 = max(1, c[0] - hw), min(kgrid.Nx, c[0] + hw)
            j0, j1 = max(1, c[1] - hh), min(kgrid.Ny, c[1] + hh)
            if i0 <= i1 and j0 <= j1:
                mask[i0 - 1:i1, j0 - 1:j1, c[2] - 1] = True
        return mask
    raise TypeError(f"no grid mapping for {type(elem).__name__}")


def get_element_binary_mask(array: KWaveArray, kgrid, element_index: int) -> np.ndarray:
    """array.jl:145-155 (1-based element_index)."""
    if not 1 <= element_index <= len(array):
        raise IndexError("element index out of range")
    return _element_mask(array.elements[element_index - 1], kgrid)


def get_array_binary_mask(array: KWaveArray, kgrid) -> np.ndarray:
    """array.jl:218-224."""
    mask = np.zeros(kgrid.grid_size(), dtype=bool)
    for i in range(1, len(array) + 1):
        mask |= get_element_binary_mask(array, kgrid, i)
    return mask


def _findall(mask: np.ndarray) -> np.ndarray:
    return np.flatnonzero(mask.ravel(order="F"))


def get_distributed_source_signal(array: KWaveArray, kgrid, source_signal) -> Tuple[np.ndarray, np.ndarray]:
    """array.jl:244-285 -- (combined mask, n_points x Nt signals): each element's
    signal divided equally over its grid points, overlapping points summed."""
    sig = np.asarray(source_signal, dtype=np.float64)
    if sig.ndim == 1:
        sig = sig[None, :]
    n = len(array)
    masks = [get_element_binary_mask(array, kgrid, i) for i in range(1, n + 1)]
    combined = np.zeros(kgrid.grid_size(), dtype=bool)
    for m in masks:
        combined |= m
    idx = _findall(combined)
    pos = np.full(combined.size, -1, dtype=np.int64)
    pos[idx] = np.arange(idx.size)
    out = np.zeros((idx.size, sig.shape[1]))
    for i, m in enumerate(masks):
        e = _findall(m)
        if e.size == 0:
            continue
        np.add.at(out, pos[e], sig[i][None, :] / e.size)
    return combined, out


def combine_sensor_data(array: KWaveArray, kgrid, sensor_data) -> np.ndarray:
    """array.jl:300-335 -- per-element mean of the rows of `sensor_data` recorded
    at the combined array mask (rows in findall order)."""
    data = np.asarray(sensor_data, dtype=np.float64)
    n = len(array)
    out = np.zeros((n, data.shape[1]))
    combined = get_array_binary_mask(array, kgrid)
    idx = _findall(combined)
    pos = np.full(combined.size, -1, dtype=np.int64)
    pos[idx] = np.arange(idx.size)
    for i in range(1, n + 1):
        rows = pos[_findall(get_element_binary_mask(array, kgrid, i))]
        rows = rows[(rows >= 0) & (rows < data.shape[0])]
        if rows.size:
            out[i - 1] = data[rows].sum(axis=0) / rows.size
    return out
