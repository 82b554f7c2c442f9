"""Exact per-Fourier-mode recurrence of the reference scheme (closed form).

TEST INFRASTRUCTURE ONLY (imported by tests/ and tests/golden/make_golden.py).

For a homogeneous medium with the PML switched off (pml_size = 0) and a
linear equation of state, every update of time_step_{1,2,3}d!
(KWave.jl/src/solver/first_order_steps.jl:296-535) is diagonal in Fourier
space.  Writing D+ = i k_d e^{+i k_d h/2}, D- = i k_d e^{-i k_d h/2}
(grid.jl:149-153, fft_utils.jl:126-182) one has D- D+ = -k_d^2 for every
mode, including the even-N Nyquist bin, so the velocity and density updates
(first_order_steps.jl:404-423) eliminate to

    rho^{n+1} - 2 rho^n + rho^{n-1} = -dt^2 |k|^2 p^n + S^{n+1} - S^n   (n >= 1)

with p^n from the equation of state (first_order_steps.jl:243-268):

    lossless        p^n = c^2 kappa rho^n
    no_dispersion   p^n = c^2 kappa (1 + nabla1) rho^n
    stokes          p^n = c^2 kappa [(1 + nabla1) rho^n - nabla2 (rho^n - prev^n)/dt]

where prev^1 = 0 and prev^n = rho^{n-1} (first_order.jl:638-639,
first_order_steps.jl:438-440), kappa = sin(c|k|dt/2)/(c|k|dt/2)
(first_order.jl:35-42) and S^n is the density injected at step n by an
additive pressure source, p_src/c^2 summed over the split components
(first_order_steps.jl:619-640).  An additive velocity source V_d^n (the raw
values added to u_d after its update at step n, first_order_steps.jl:404-415,
541-559, 586-617, 646-690) enters through the density update of the same
step, rho_d^n = rho_d^{n-1} - dt rho0 D-_d u_d^n, so the recurrence gains

    - dt rho0 sum_d D-_d V_d^{n+1}

on its right-hand side (and rho^1 gains - dt rho0 sum_d D-_d V_d^1).  The
start is rho^0 = 0, rho^1 = S^1 for a source-driven run, and rho^0 = p0/c^2, p^0 = p0, u^0 = -dt/(2 rho0) D+ p0,
giving rho^1 = p0/c^2 - 1.5 dt^2 |k|^2 p0 + S^1 for an initial-value run
(first_order_steps.jl:97-137; smooth_p0 = false).

Nothing here uses the oracle's (or the product's) FFT plumbing: the input
and output transforms are full complex numpy.fft.fftn/ifftn, and the time
stepping is a scalar recurrence per mode.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np


def _kvec(N: int, d: float) -> np.ndarray:
    dk = 2 * math.pi / (N * d)
    return np.fft.fftfreq(N, 1.0 / N) * dk   # integers [0..N/2-1, -N/2..-1] (odd: symmetric)


def _k2(shape, d):
    k2 = np.zeros(shape)
    for a, (n, h) in enumerate(zip(shape, d)):
        s = [1] * len(shape)
        s[a] = n
        k2 = k2 + _kvec(n, h).reshape(s) ** 2
    return k2


def recurrence_trace(shape: Sequence[int], d: Sequence[float], c0: float, rho0: float,
                     dt: float, Nt: int, sensor_idx: np.ndarray,
                     src_idx: Optional[np.ndarray] = None, src_signal: Optional[np.ndarray] = None,
                     p0: Optional[np.ndarray] = None, alpha: Optional[dict] = None,
                     c_ref: Optional[float] = None, usrc_idx: Optional[np.ndarray] = None,
                     usrc_signals: Optional[Sequence[Optional[np.ndarray]]] = None):
    """Sensor pressure traces (n_sensor x Nt) and the final pressure field.

    ``src_idx`` are column-major linear indices, ``src_signal`` is
    (n_src x ncols) sampled at min(t, ncols) (first_order_steps.jl:625).
    ``alpha`` = dict(tau, eta, y, mode) overrides the absorption prefactors.
    ``usrc_idx`` / ``usrc_signals`` (one (n x ncols) array or None per axis)
    describe an additive velocity source.
    """
    shape = tuple(int(n) for n in shape)
    nd = len(shape)
    c_ref = c0 if c_ref is None else c_ref
    k2 = _k2(shape, d)
    k = np.sqrt(k2)
    arg = c_ref * k * dt / 2
    kappa = np.ones_like(arg)
    nz = arg != 0
    kappa[nz] = np.sin(arg[nz]) / arg[nz]
    n1 = np.zeros_like(k)
    n2 = np.zeros_like(k)
    mode = "lossless"
    if alpha is not None:
        mode = alpha["mode"]
        with np.errstate(divide="ignore", invalid="ignore"):
            n1 = alpha["tau"] * k ** alpha["y"]
            n2 = alpha["eta"] * k ** (alpha["y"] - 1)
    c2 = c0 * c0

    def eos(rho_n, prev):
        if mode == "lossless":
            return c2 * kappa * rho_n
        if mode == "no_dispersion":
            return c2 * kappa * (1 + n1) * rho_n
        return c2 * kappa * ((1 + n1) * rho_n - n2 * (rho_n - prev) / dt)

    def src_hat(t):
        if src_idx is None:
            return 0.0
        col = min(t, src_signal.shape[1]) - 1
        f = np.zeros(int(np.prod(shape)))
        np.add.at(f, src_idx, src_signal[:, col] / c2)
        return np.fft.fftn(f.reshape(shape, order="F"))

    # staggered divergence multipliers D-_d = i k_d exp(-i k_d h/2) on the full grid
    dminus = []
    for a, (n, h) in enumerate(zip(shape, d)):
        kv = _kvec(n, h)
        sh = [1] * nd
        sh[a] = n
        dminus.append((1j * kv * np.exp(-1j * kv * h / 2)).reshape(sh))

    def vel_hat(t):
        """- dt rho0 sum_d D-_d FFT(V_d^t), the density forcing of the velocity source at step t."""
        if usrc_idx is None or usrc_signals is None:
            return 0.0
        acc = np.zeros(shape, dtype=complex)
        for a, sig in enumerate(usrc_signals):
            if sig is None:
                continue
            col = min(t, sig.shape[1]) - 1
            f = np.zeros(int(np.prod(shape)))
            np.add.at(f, usrc_idx, sig[:, col])
            acc += dminus[a] * np.fft.fftn(f.reshape(shape, order="F"))
        return -dt * rho0 * acc

    sens = np.asarray(sensor_idx)
    out = np.zeros((sens.size, Nt))
    zero = np.zeros(shape, dtype=complex)
    if p0 is not None:
        P0 = np.fft.fftn(np.asarray(p0, dtype=np.float64))
        rho_prev_state = P0 / c2                                   # rho^0
        rho_n = P0 / c2 - 1.5 * dt * dt * k2 * P0 + src_hat(1) + vel_hat(1)   # rho^1
    else:
        rho_prev_state = zero
        rho_n = src_hat(1) + vel_hat(1) + zero
    S_prev = src_hat(1)
    eos_prev = zero                                                # prev^1 = 0
    p_hat = None
    for t in range(1, Nt + 1):
        p_hat = eos(rho_n, eos_prev)
        p = np.fft.ifftn(p_hat).real
        out[:, t - 1] = p.ravel(order="F")[sens]
        if t == Nt:
            break
        S_next = src_hat(t + 1)
        rho_next = 2 * rho_n - rho_prev_state - dt * dt * k2 * p_hat + (S_next - S_prev) + vel_hat(t + 1)
        rho_prev_state, eos_prev = rho_n, rho_n
        rho_n, S_prev = rho_next, S_next
    return out, np.fft.ifftn(p_hat).real
