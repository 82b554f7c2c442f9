"""Host-side input builders the configurations use (no GPU work here).

tone_burst (signal/generation.jl:30-71), make_disc / make_circle
(geometry/shapes.jl:22-107), make_arc (130-168), make_ball (260-278),
make_bowl (330-377).
Julia's ``round(Int, x)`` rounds ties to even, like Python's ``round``.
"""
from __future__ import annotations

import math

import numpy as np


def tone_burst(sample_freq: float, signal_freq: float, num_cycles: int, envelope: str = "gaussian",
               signal_offset: int = 0) -> np.ndarray:
    """generation.jl:30-71."""
    tone_length = num_cycles / signal_freq
    dt = 1.0 / sample_freq
    n = int(round(tone_length / dt))
    t = np.arange(n) * dt
    carrier = np.sin(2 * math.pi * signal_freq * t)
    if envelope == "gaussian":
        sigma = tone_length / (2 * num_cycles)
        env = np.exp(-((t - tone_length / 2) / sigma) ** 2 / 2)
    elif envelope == "rectangular":
        env = np.ones(n)
    elif envelope == "cosine":
        env = 0.5 * (1.0 - np.cos(2 * math.pi * np.arange(n) / (n - 1)))
    else:
        raise ValueError(f"Unknown envelope type: {envelope}. Use :gaussian, :rectangular, or :cosine")
    sig = carrier * env
    if signal_offset > 0:
        sig = np.concatenate([np.zeros(signal_offset), sig])
    return sig


def make_disc(Nx: int, Ny: int, cx, cy, radius) -> np.ndarray:
    """shapes.jl:22-38 (1-based centre)."""
    cx, cy = int(round(cx)), int(round(cy))
    i = np.arange(1, Nx + 1)[:, None]
    j = np.arange(1, Ny + 1)[None, :]
    return (i - cx) ** 2 + (j - cy) ** 2 <= radius ** 2


def make_circle(Nx: int, Ny: int, cx, cy, radius) -> np.ndarray:
    """shapes.jl:56-107 -- full circle rasterised by the midpoint algorithm."""
    cx, cy, r = int(round(cx)), int(round(cy)), int(round(radius))
    pts = []
    x, y, d = r, 0, 1 - r
    while x >= y:
        pts += [(cx + x, cy + y), (cx - x, cy + y), (cx + x, cy - y), (cx - x, cy - y),
                (cx + y, cy + x), (cx - y, cy + x), (cx + y, cy - x), (cx - y, cy - x)]
        y += 1
        if d <= 0:
            d += 2 * y + 1
        else:
            x -= 1
            d += 2 * (y - x) + 1
    mask = np.zeros((Nx, Ny), dtype=bool)
    for px, py in pts:
        if 1 <= px <= Nx and 1 <= py <= Ny:
            mask[px - 1, py - 1] = True
    return mask


def make_arc(Nx: int, Ny: int, cx, cy, radius, diameter, focus) -> np.ndarray:
    """shapes.jl:130-168 -- single-pixel arc sampled at max(200, ...) angles around
    the centre of curvature (cx, cy), opening towards `focus` (1-based indices)."""
    cx, cy = int(round(cx)), int(round(cy))
    fx, fy = int(round(focus[0])), int(round(focus[1]))
    arc = np.zeros((Nx, Ny), dtype=bool)
    dx, dy = fx - cx, fy - cy
    if math.isclose(math.sqrt(dx * dx + dy * dy), 0.0, abs_tol=1e-12):
        return arc
    center_angle = math.atan2(dy, dx)
    half = math.asin(min(max(diameter / (2 * radius), -1.0), 1.0))
    n_points = max(200, int(math.ceil(4 * radius * 2 * half / (2 * math.pi))))
    for i in range(n_points + 1):
        th = center_angle - half + i * 2 * half / n_points
        px = int(np.rint(cx + radius * math.cos(th)))
        py = int(np.rint(cy + radius * math.sin(th)))
        if 1 <= px <= Nx and 1 <= py <= Ny:
            arc[px - 1, py - 1] = True
    return arc


def make_ball(Nx: int, Ny: int, Nz: int, cx, cy, cz, radius) -> np.ndarray:
    """shapes.jl:260-278."""
    cx, cy, cz = int(round(cx)), int(round(cy)), int(round(cz))
    i = np.arange(1, Nx + 1)[:, None, None]
    j = np.arange(1, Ny + 1)[None, :, None]
    k = np.arange(1, Nz + 1)[None, None, :]
    return (i - cx) ** 2 + (j - cy) ** 2 + (k - cz) ** 2 <= radius ** 2


def make_bowl(Nx: int, Ny: int, Nz: int, center, radius, diameter, focus) -> np.ndarray:
    """shapes.jl:330-377; ``center`` is the centre of curvature, ``focus`` only orients."""
    cx, cy, cz = (int(round(v)) for v in center)
    fx, fy, fz = (int(round(v)) for v in focus)
    dvec = np.array([fx - cx, fy - cy, fz - cz], dtype=np.float64)
    nrm = float(np.sqrt(np.sum(dvec ** 2)))
    bowl = np.zeros((Nx, Ny, Nz), dtype=bool)
    if nrm == 0.0:
        return bowl
    dvec /= nrm
    cos_half = math.cos(math.asin(min(max(diameter / (2 * radius), -1.0), 1.0)))
    r2i, r2o = (radius - 0.5) ** 2, (radius + 0.5) ** 2
    j = (np.arange(1, Ny + 1) - cy)[:, None]
    k = (np.arange(1, Nz + 1) - cz)[None, :]
    for ii in range(Nx):          # slab by slab keeps a 512^3 build within memory
        i = ii + 1 - cx
        d2 = i * i + j * j + k * k
        shell = (d2 >= r2i) & (d2 <= r2o)
        if not shell.any():
            continue
        with np.errstate(invalid="ignore", divide="ignore"):
            cosang = (i * dvec[0] + j * dvec[1] + k * dvec[2]) / np.sqrt(d2)
        bowl[ii] = shell & (cosang >= cos_half)
    return bowl
