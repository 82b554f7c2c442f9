"""Write the golden fixtures under tests/golden/ (run from the repo root).

The fixtures come from the exact per-mode recurrence of the reference scheme
(oracle/closed_form.py: homogeneous medium, pml_size = 0, linear EOS), which
uses no part of the oracle's or the product's FFT plumbing, plus the
128^2 example inputs (tone burst, disc and circle masks) that the
reference's own tests constrain (test_signal.jl, test_geometry.jl).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.closed_form import recurrence_trace  # noqa: E402


def _mask(rng, shape, n):
    idx = np.sort(rng.choice(int(np.prod(shape)), size=n, replace=False))
    return idx.astype(np.int64)


def db2neper(a, y):
    # conversion.jl:34 (restated here so the fixture does not depend on the oracle module)
    return a * (1e-2 / (20 * math.log10(math.e))) * (2 * math.pi * 1e6) ** (-y)


def recurrence_case(name, shape, nt, mode=None, alpha=0.0, y=1.5, p0=False, seed=1):
    rng = np.random.default_rng(seed)
    d = 1e-4
    c0, rho0 = 1500.0, 1000.0
    dt = 0.3 * d / c0
    nd = len(shape)
    nvox = int(np.prod(shape))
    src_idx = _mask(rng, shape, max(2, nvox // 300))
    sens_idx = _mask(rng, shape, max(4, nvox // 100))
    ncols = nt // 2
    sig = rng.standard_normal((src_idx.size, ncols)) * np.hanning(ncols)[None, :]
    P0 = None
    if p0:
        P0 = np.zeros(shape)
        P0[tuple(slice(n // 2 - 2, n // 2 + 2) for n in shape)] = 1.0
        P0 += 0.05 * rng.standard_normal(shape)
    ab = None
    if mode is not None:
        a_np = db2neper(alpha, y)
        ab = dict(tau=-2 * a_np * c0 ** (y - 1), eta=2 * a_np * c0 ** (y - 1) * math.tan(math.pi * y / 2),
                  y=y, mode=mode)
    traces, pfin = recurrence_trace(shape, [d] * nd, c0, rho0, dt, nt, sens_idx, src_idx, sig, P0, ab)
    out = dict(shape=np.array(shape), d=d, c0=c0, rho0=rho0, dt=dt, Nt=nt, src_idx=src_idx, src_signal=sig,
               sensor_idx=sens_idx, p=traces, p_final_field=pfin.ravel(order="F"),
               alpha_coeff=alpha, alpha_power=y, alpha_mode="" if mode is None else mode)
    if P0 is not None:
        out["p0"] = P0.ravel(order="F")
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, shape, nt, "max|p|", float(np.abs(traces).max()))


def c1_inputs():
    """Inputs of the 128^2 example: KWaveGrid(128,1e-4,128,1e-4), make_time!(kgrid,1500)."""
    dt = 0.3 * 1e-4 / 1500.0
    fs = 1 / dt
    n = int(round((5 / 1e6) / dt))
    t = np.arange(n) * dt
    tl = 5 / 1e6
    burst = np.sin(2 * math.pi * 1e6 * t) * np.exp(-((t - tl / 2) / (tl / 10)) ** 2 / 2)
    # disc radius 5 at (64, 64), 1-based: brute force over the grid
    disc = [(i, j) for j in range(1, 129) for i in range(1, 129) if (i - 64) ** 2 + (j - 64) ** 2 <= 25]
    disc_idx = np.array([(i - 1) + 128 * (j - 1) for i, j in disc], dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "c1_inputs.npz"), dt=dt, fs=fs, tone_burst=burst,
                        disc_idx=np.sort(disc_idx))
    print("wrote c1_inputs: burst", burst.size, "disc", disc_idx.size)


if __name__ == "__main__":
    recurrence_case("rec2d_lossless", (64, 48), 200)
    recurrence_case("rec2d_p0", (48, 64), 150, p0=True, seed=2)
    recurrence_case("rec2d_nodisp", (64, 64), 150, mode="no_dispersion", alpha=5e3, seed=3)
    recurrence_case("rec2d_stokes", (48, 48), 150, mode="stokes", alpha=2.0, seed=4)
    recurrence_case("rec3d_lossless", (32, 24, 16), 100, seed=5)
    recurrence_case("rec3d_stokes_p0", (16, 24, 32), 80, mode="stokes", alpha=2.0, p0=True, seed=6)
    recurrence_case("rec1d_p0", (256,), 300, p0=True, seed=7)
    c1_inputs()
