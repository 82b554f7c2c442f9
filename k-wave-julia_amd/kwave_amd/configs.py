"""Synthetic inputs of the benchmark configurations (seeded, reproducible).

A splitmix64 stream (seed 42) drives every random draw so the inputs are the
same from any host language.  Only host-side array construction lives here.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np

from .builders import make_bowl, make_circle, make_disc, tone_burst
from .model import KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, make_time


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014); uniform() uses the top 53 bits."""

    def __init__(self, seed: int = 42):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def homogeneous_2d(data_cast=np.float64):
    """2-D 128x128, dx = 0.1 mm, homogeneous lossless; disc p source, circle sensor."""
    kg = KWaveGrid(128, 1e-4, 128, 1e-4)
    make_time(kg, 1500.0)
    med = KWaveMedium(sound_speed=1500.0, density=1000.0)
    dt, Nt = kg.dt, kg.Nt
    sig = tone_burst(1 / dt, 1e6, 5)
    sig = np.concatenate([sig, np.zeros(max(0, Nt - sig.size))])[:Nt]
    pm = make_disc(128, 128, 64, 64, 5)
    src = KWaveSource(p_mask=pm, p=np.repeat(sig[None, :], int(pm.sum()), axis=0))
    sen = KWaveSensor(mask=make_circle(128, 128, 64, 64, 40), record=["p", "p_final"])
    return kg, med, src, sen, dict(pml_size=20, pml_alpha=2.0, data_cast=data_cast)


def ellipsoids_3d(N: int = 256, Nt: int = 1200, pml: int = 20, amplitude: float = 1e6, seed: int = 42):
    """3-D heterogeneous absorbing nonlinear FP32 workload (N^3, dx = 0.1 mm).

    c0 = 1500 m/s plus 8 axis-aligned ellipsoids (centres in [pml+1, N-pml]^3,
    integer semi-axes in [10, 40], c0 ~ U[1400,1650], rho0 ~ U[900,1100];
    later ones overwrite earlier ones); alpha0 = 0.75, y = 1.5, :stokes
    (heterogeneous c0 -> absorption without dispersion), B/A = 5; plane p
    source at z = pml+1 driven by a 1 MHz 3-cycle Gaussian tone burst of
    `amplitude` Pa; sensor plane z = N - 56 recording p, p_max, p_rms.
    The signal matrix holds the burst plus one zero column: past the last
    column the reference re-adds the last value (first_order_steps.jl:625),
    so this equals zero padding to Nt.
    """
    rng = SplitMix64(seed)
    kg = KWaveGrid(N, 1e-4, N, 1e-4, N, 1e-4)
    c0 = np.full((N, N, N), 1500.0)
    rho0 = np.full((N, N, N), 1000.0)
    lo, hi = pml + 1, N - pml
    x = np.arange(1, N + 1)
    for _ in range(8):
        cen = [rng.uniform(lo, hi) for _ in range(3)]
        ax = [rng.integer(10, 40) for _ in range(3)]
        cv = rng.uniform(1400.0, 1650.0)
        rv = rng.uniform(900.0, 1100.0)
        sl = [slice(max(0, int(math.floor(c - a)) - 1), min(N, int(math.ceil(c + a)) + 1)) for c, a in zip(cen, ax)]
        X = ((x[sl[0]] - cen[0]) / ax[0])[:, None, None] ** 2
        Y = ((x[sl[1]] - cen[1]) / ax[1])[None, :, None] ** 2
        Z = ((x[sl[2]] - cen[2]) / ax[2])[None, None, :] ** 2
        inside = (X + Y + Z) <= 1.0
        c0[sl[0], sl[1], sl[2]][inside] = cv
        rho0[sl[0], sl[1], sl[2]][inside] = rv
    make_time(kg, c0)
    kg.Nt = Nt
    kg.t_array = np.arange(Nt) * kg.dt
    med = KWaveMedium(sound_speed=c0, density=rho0, alpha_coeff=0.75, alpha_power=1.5, alpha_mode="stokes",
                      BonA=5.0)
    burst = amplitude * tone_burst(1 / kg.dt, 1e6, 3)
    pm = np.zeros((N, N, N), dtype=bool)
    pm[:, :, pml] = True                                   # z = pml + 1 (1-based)
    n_src = int(pm.sum())
    sig = np.zeros((n_src, burst.size + 1))
    sig[:, : burst.size] = burst[None, :]
    src = KWaveSource(p_mask=pm, p=sig)
    sm = np.zeros((N, N, N), dtype=bool)
    sm[:, :, N - 57] = True                                # z = 200 for N = 256 (1-based)
    sen = KWaveSensor(mask=sm, record=["p", "p_max", "p_rms"])
    return kg, med, src, sen, dict(pml_size=pml, pml_alpha=2.0, data_cast=np.float32, smooth_p0=False)


def ellipsoids_3d_bytes_per_voxel_update(precision: int = 4) -> int:
    """Compulsory bytes per voxel-update: 7 evolving fields read+written, 2 maps read."""
    return 2 * 7 * precision + 2 * precision


def uneven_2d_batch(n_sims: int = 64, Nt: int = 1500, seed: int = 42, sizes=(256, 384, 512, 768, 1024)):
    """Batch of 2-D firings of a 128-element linear array (pitch 3, width 2 cells).

    Per simulation: an N x N grid with N drawn (seeded) from `sizes`, dx = 0.1 mm;
    c0 = 1525 + 25 z clipped to [1450, 1600] m/s where z is a standardised
    Gaussian random field (white noise smoothed with a periodic Gaussian of
    sigma = 10 cells = 1 mm); rho0 = 1000; alpha0 = 0.5, y = 1.1, :stokes;
    the array sits on row x = 21 (1-based) centred in y, elements that do not
    fit the grid are dropped; simulation i fires element i mod (elements kept)
    with a 2.5 MHz 3-cycle tone burst; every kept array cell records p.
    Returns a list of (kgrid, medium, source, sensor) and the shared kwargs.
    """
    from scipy import ndimage

    rng = SplitMix64(seed)
    sims = []
    for i in range(n_sims):
        N = sizes[rng.integer(0, len(sizes) - 1)]
        noise = np.random.default_rng(seed + 1000 + i).standard_normal((N, N))
        z = ndimage.gaussian_filter(noise, sigma=10.0, mode="wrap")
        z = (z - z.mean()) / z.std()
        c0 = np.clip(1525.0 + 25.0 * z, 1450.0, 1600.0)
        kg = KWaveGrid(N, 1e-4, N, 1e-4)
        make_time(kg, c0)
        kg.Nt = Nt
        kg.t_array = np.arange(Nt) * kg.dt
        med = KWaveMedium(sound_speed=c0, density=1000.0, alpha_coeff=0.5, alpha_power=1.1, alpha_mode="stokes")
        span = 127 * 3 + 2
        y0 = (N - span) // 2
        elems = []
        for e in range(128):
            ys = [y0 + 3 * e, y0 + 3 * e + 1]
            if ys[0] >= 0 and ys[1] < N:
                elems.append(ys)
        fire = elems[i % len(elems)]
        pm = np.zeros((N, N), dtype=bool)
        pm[20, fire] = True
        burst = tone_burst(1 / kg.dt, 2.5e6, 3)
        sig = np.zeros((2, burst.size + 1))
        sig[:, : burst.size] = burst[None, :]
        src = KWaveSource(p_mask=pm, p=sig)
        sm = np.zeros((N, N), dtype=bool)
        for ys in elems:
            sm[20, ys] = True
        sen = KWaveSensor(mask=sm, record=["p"])
        sims.append((kg, med, src, sen))
    return sims, dict(pml_size=20, pml_alpha=2.0, smooth_p0=False)


def layered_3d_fp16(N: int = 512, Nt: int = 2000, pml: int = 24, amplitude: float = 1e6, seed: int = 42,
            n_points: int = 4096):
    """3-D layered FP16-map workload (N^3, dx = 0.2 mm, PML 24).

    Six layers split at 5 seeded z-planes (uniform in [pml+1, N-pml], sorted),
    c0 ~ U[1400,1700] and rho0 ~ U[950,1150] per layer, plus N(0, 10 m/s)
    noise on c0 (numpy default_rng(seed)); B/A = 6, alpha0 = 0.6, y = 1.5,
    :stokes with scalar alpha (an array alpha fails in the reference,
    first_order.jl:588-592).  Bowl source make_bowl(N,N,N,(N/2,N/2,175),150,120,
    (N/2,N/2,0)) (1-based), a 1 MHz 10-cycle tone burst of `amplitude` Pa.
    Two sensor sets (one KWaveSensor cannot hold both, sensor.jl:31-38): the
    whole volume recording p_max, and `n_points` seeded points recording p.
    Returns (kgrid, medium, source, [sensor_pmax, sensor_points], kwargs).
    """
    rng = SplitMix64(seed)
    kg = KWaveGrid(N, 2e-4, N, 2e-4, N, 2e-4)
    cuts = sorted(rng.integer(pml + 1, N - pml) for _ in range(5))
    c_lay = [rng.uniform(1400.0, 1700.0) for _ in range(6)]
    r_lay = [rng.uniform(950.0, 1150.0) for _ in range(6)]
    z = np.arange(1, N + 1)
    layer = np.searchsorted(np.asarray(cuts), z, side="right")          # 0..5 along z
    c0 = np.broadcast_to(np.asarray(c_lay)[layer][None, None, :], (N, N, N)).astype(np.float64)
    noise = np.random.default_rng(seed).standard_normal((N, N, N), dtype=np.float32)
    c0 = c0 + 10.0 * noise
    del noise
    rho0 = np.broadcast_to(np.asarray(r_lay)[layer][None, None, :], (N, N, N)).astype(np.float64)
    make_time(kg, c0)
    kg.Nt = Nt
    kg.t_array = np.arange(Nt) * kg.dt
    med = KWaveMedium(sound_speed=c0, density=rho0, alpha_coeff=0.6, alpha_power=1.5, alpha_mode="stokes",
                      BonA=6.0)
    h = N // 2
    pm = make_bowl(N, N, N, (h, h, 175), 150, 120, (h, h, 0))
    burst = amplitude * tone_burst(1 / kg.dt, 1e6, 10)
    n_src = int(pm.sum())
    sig = np.zeros((n_src, burst.size + 1))
    sig[:, : burst.size] = burst[None, :]
    src = KWaveSource(p_mask=pm, p=sig)
    s_all = KWaveSensor(mask=np.ones((N, N, N), dtype=bool), record=["p_max"])
    pts = np.zeros(N ** 3, dtype=bool)
    pts[np.random.default_rng(seed + 1).choice(N ** 3, size=n_points, replace=False)] = True
    s_pts = KWaveSensor(mask=pts.reshape((N, N, N), order="F"), record=["p"])
    return kg, med, src, [s_all, s_pts], dict(pml_size=pml, pml_alpha=2.0, smooth_p0=False,
                                              medium_storage=np.float16)


def ring_2d_adjoint(N: int = 512, Nt: int = 1600, n_elements: int = 128, radius: float = 220.0, seed: int = 42):
    """Adjoint workload: 2-D N x N, dx = 0.1 mm, rho0 = 1000, lossless linear.

    True model: c0 = 1500 plus 6 seeded discs (radius integer in [10,40],
    c0 ~ U[1450,1600], centres uniform in the ring's interior); starting
    model: homogeneous 1500 given as a map.  dt comes from the true model
    (CFL 0.3 on its maximum) and is held fixed for every model.  Ring of
    `n_elements` points at round(N/2 + radius (cos, sin)(2 pi k / n)) (1-based);
    firing k drives element k with a 1.5 MHz 2-cycle tone burst; every
    element records p.  Returns (kgrid, c_true, c_start, firings, kwargs) where
    firings is a list of (source, sensor).
    """
    rng = SplitMix64(seed)
    kg = KWaveGrid(N, 1e-4, N, 1e-4)
    c_true = np.full((N, N), 1500.0)
    h = N / 2
    for _ in range(6):
        r = rng.integer(10, 40)
        cx = rng.uniform(h - radius / 2, h + radius / 2)
        cy = rng.uniform(h - radius / 2, h + radius / 2)
        cv = rng.uniform(1450.0, 1600.0)
        c_true[make_disc(N, N, int(round(cx)), int(round(cy)), r)] = cv
    make_time(kg, c_true)
    kg.Nt = Nt
    kg.t_array = np.arange(Nt) * kg.dt
    ang = 2 * np.pi * np.arange(n_elements) / n_elements
    ex = np.round(h + radius * np.cos(ang)).astype(int) - 1          # 0-based
    ey = np.round(h + radius * np.sin(ang)).astype(int) - 1
    ring = np.zeros((N, N), dtype=bool)
    ring[ex, ey] = True
    burst = tone_burst(1 / kg.dt, 1.5e6, 2)
    firings = []
    for k in range(n_elements):
        pm = np.zeros((N, N), dtype=bool)
        pm[ex[k], ey[k]] = True
        sig = np.zeros((1, burst.size + 1))
        sig[0, : burst.size] = burst
        firings.append((KWaveSource(p_mask=pm, p=sig), KWaveSensor(mask=ring, record=["p"])))
    c_start = np.full((N, N), 1500.0)
    return kg, c_true, c_start, firings, dict(pml_size=20, pml_alpha=2.0, smooth_p0=False)


def two_layer_elastic_3d(N: int = 256, Nt: int = 600, pml: int = 10, seed: int = 42):
    """3-D two-layer elastic workload (N^3, dx = 0.1 mm, PML 10) for pstd_elastic_3d.

    z < N/2: fluid-like layer (cp 1540 m/s, cs 0, rho 1000); z >= N/2: solid
    (cp 3000, cs 1500, rho 1900); N(0, 10 m/s) seeded noise on cp.  Initial
    pressure: a ball of radius 6 cells, 1 MPa, centred at (N/2, N/2, N/4);
    sensor: the plane z = N/2 - 8 recording p and p_max.
    Returns (kgrid, medium, source, sensor, kwargs).
    """
    from .builders import make_ball
    from .model import ElasticMedium, ElasticSource

    kg = KWaveGrid(N, 1e-4, N, 1e-4, N, 1e-4)
    solid = (np.arange(N) >= N // 2)[None, None, :]
    noise = np.random.default_rng(seed).standard_normal((N, N, N), dtype=np.float32)
    cp = np.where(solid, 3000.0, 1540.0) + 10.0 * noise
    del noise
    cs = np.broadcast_to(np.where(solid, 1500.0, 0.0), (N, N, N)).astype(np.float64)
    rho = np.broadcast_to(np.where(solid, 1900.0, 1000.0), (N, N, N)).astype(np.float64)
    make_time(kg, cp)
    kg.Nt = Nt
    kg.t_array = np.arange(Nt) * kg.dt
    med = ElasticMedium(sound_speed_compression=cp, sound_speed_shear=cs, density=rho)
    src = ElasticSource(p0=1e6 * make_ball(N, N, N, N // 2, N // 2, N // 4, 6).astype(np.float64))
    mask = np.zeros((N, N, N), dtype=bool)
    mask[:, :, N // 2 - 8] = True
    sen = KWaveSensor(mask=mask, record=["p", "p_max"])
    return kg, med, src, sen, {"pml_size": pml, "pml_alpha": 2.0}


def elastic_pass_bytes(precision: int = 4):
    """Bytes per voxel each class of elastic pass must move on the 3-D heterogeneous
    workload, as (bytes per launch summed over a step, launches per step).

    A derivative pass reads its input and writes (2 fields) or accumulates into
    (3 fields) a scratch field; a velocity pass reads the stress, the partial
    sum, v and rho and writes v (5); the normal-stress pass reads v_z, two
    derivatives, lambda, mu and three stresses and writes three (11); a shear
    pass reads v, the partial sum, mu and the stress and writes it (5).
    """
    s = precision
    return {
        "k_el_deriv_x": (3 * 2 * s, 3), "k_el_deriv_y": ((3 * 2 + 3) * s, 4),
        "k_el_deriv_z": ((2 * 3 + 2 * 2) * s, 4), "k_el_vel_x": (5 * s, 1), "k_el_vel_y": (5 * s, 1),
        "k_el_vel_z": (5 * s, 1), "k_el_normal_z": (11 * s, 1), "k_el_shear_x": (2 * 5 * s, 2),
        "k_el_shear_y": (5 * s, 1),
    }
