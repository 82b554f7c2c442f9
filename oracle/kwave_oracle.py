"""FP64 CPU restatement of KWave.jl's k-space first-order PSTD time loop.

TEST INFRASTRUCTURE ONLY.  This module is the oracle: it is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py`` -- as the parity reference and the CPU baseline, never as the product.
The product path (``kwave_amd`` + the HIP library) must not import it.

Every function names the reference file:line it restates (paths relative to
the reference repository root, ``KWave.jl/src/...``).  Arrays follow the
reference's column-major convention expressed in NumPy: a field of grid
shape ``(Nx, Ny, Nz)`` is indexed ``[i, j, k]`` (0-based) and the reference's
linear index ``findall`` order is NumPy's Fortran order.

Parity status of each part (what pins it; see tests/test_oracle_*.py):
  grid vectors / make_time ........ pinned: test_grid.jl assertions + closed forms
  get_pml ......................... pinned: test_pml.jl assertions + closed forms
  db2neper, kappa, absorption ..... pinned: closed forms (conversion.jl:34)
  tone_burst, masks ............... pinned: test_signal.jl / test_geometry.jl + brute force
  spectral_gradient ............... pinned: analytic sinusoids (test_fft.jl:24-56),
                                    staggered shift, Nyquist, dense DFT matrices
  time loop, lossless / absorbing . pinned: exact per-mode recurrence (closed form)
  p0 initialisation ............... pinned: exact per-mode recurrence
  heterogeneous c0/rho0 ........... pinned: uniform-array == scalar branch identity,
                                    dense-matrix brute force on tiny grids
  B/A nonlinearity ................ pinned: dense-matrix brute force on tiny grids
                                    (same update read from first_order_steps.jl)
  time reversal ................... pinned: TR peak invariant (test_time_reversal.jl)
  smooth .......................... pinned: window closed form + restore_max invariant
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence

import numpy as np
import scipy.fft as sfft

DIRICHLET, ADDITIVE, ADDITIVE_NO_CORRECTION = 0, 1, 2      # types.jl:14

VALID_RECORD_FIELDS = {                                       # sensor.jl:41-45
    "p", "p_max", "p_min", "p_rms", "p_final",
    "ux", "uy", "uz", "u_max", "u_rms", "u_final",
    "I_avg", "I_max",
}

_WORKERS = 1


def set_fft_workers(n: int) -> None:
    """Threads used by scipy.fft (the FFTW stand-in, fft_utils.jl:53-55)."""
    global _WORKERS
    _WORKERS = max(1, int(n))


# ---------------------------------------------------------------------------
# Grid (grid.jl)
# ---------------------------------------------------------------------------

def make_wavenumber_vec(N: int, d: float) -> np.ndarray:
    """grid.jl:120-130 -- [0..N/2-1, -N/2..-1]*2pi/(N d) (odd N: symmetric)."""
    dk = 2 * math.pi / (N * d)
    if N % 2 == 0:
        idx = np.concatenate([np.arange(0, N // 2), np.arange(-(N // 2), 0)])
    else:
        idx = np.concatenate([np.arange(0, (N - 1) // 2 + 1), np.arange(-((N - 1) // 2), 0)])
    return idx.astype(np.float64) * dk


def make_spatial_vec(N: int, d: float) -> np.ndarray:
    """grid.jl:137-139."""
    return (np.arange(N) - (N - 1) / 2.0) * d


def make_shift_operators(kv: np.ndarray, d: float):
    """grid.jl:149-153 -- exp(+-i k d/2)."""
    return np.exp(1j * kv * d / 2), np.exp(-1j * kv * d / 2)


@dataclass
class Grid:
    """Restatement of KWaveGrid1D/2D/3D (grid.jl:21-108, ctors 183-245)."""
    N: tuple
    d: tuple
    dt: float = 0.0
    Nt: int = 0

    def __post_init__(self):
        self.N = tuple(int(n) for n in self.N)
        self.d = tuple(float(x) for x in self.d)
        self.k_vecs = [make_wavenumber_vec(n, dd) for n, dd in zip(self.N, self.d)]
        self.shifts = [make_shift_operators(kv, dd) for kv, dd in zip(self.k_vecs, self.d)]
        # |k| on the full grid, summed in the reference's order (grid.jl:204, 228-232)
        if len(self.N) == 1:
            self.k = np.abs(self.k_vecs[0])
        elif len(self.N) == 2:
            self.k = np.sqrt(self.k_vecs[0][:, None] ** 2 + self.k_vecs[1][None, :] ** 2)
        else:
            self.k = np.sqrt(self.k_vecs[0][:, None, None] ** 2
                             + self.k_vecs[1][None, :, None] ** 2
                             + self.k_vecs[2][None, None, :] ** 2)

    @property
    def ndim(self) -> int:
        return len(self.N)

    @property
    def t_array(self) -> np.ndarray:
        return np.arange(self.Nt) * self.dt


def grid_from(kgrid: Any) -> Grid:
    """Build the oracle grid from any object carrying Nx/dx[/Ny/dy[/Nz/dz]], dt, Nt."""
    N = [kgrid.Nx]
    d = [kgrid.dx]
    if hasattr(kgrid, "Ny") and getattr(kgrid, "Ny", None) is not None:
        N.append(kgrid.Ny)
        d.append(kgrid.dy)
    if hasattr(kgrid, "Nz") and getattr(kgrid, "Nz", None) is not None:
        N.append(kgrid.Nz)
        d.append(kgrid.dz)
    dt = kgrid.dt if not callable(kgrid.dt) else kgrid.dt()
    Nt = kgrid.Nt if not callable(kgrid.Nt) else kgrid.Nt()
    return Grid(tuple(N), tuple(d), float(dt), int(Nt))


def make_time(N: Sequence[int], d: Sequence[float], sound_speed, cfl: float = 0.3,
              t_end: Optional[float] = None):
    """grid.jl:320-352 -- dt = cfl*min(d)/max(c); Nt = ceil(t_end/dt)+1."""
    c = np.asarray(sound_speed, dtype=np.float64)
    c_max, c_min = float(np.max(c)), float(np.min(c))
    d_min = min(d)
    dt = cfl * d_min / c_max
    if t_end is None:
        diag = math.sqrt(sum((n * dd) ** 2 for n, dd in zip(N, d)))
        t_end = 2.0 * diag / c_min
    Nt = int(math.ceil(t_end / dt)) + 1
    return dt, Nt


# ---------------------------------------------------------------------------
# PML, units, kappa, absorption (pml.jl, conversion.jl, first_order.jl)
# ---------------------------------------------------------------------------

def get_pml(N: int, d: float, pml_size: int, pml_alpha: float, dt: float,
            staggered: bool = False) -> np.ndarray:
    """pml.jl:25-55 -- exp(-alpha*pos^4*dt) on both ends, 1 inside."""
    pml = np.ones(N, dtype=np.float64)
    if pml_size == 0:
        return pml
    for i in range(1, pml_size + 1):
        if staggered:
            pos = (pml_size - i + 0.5) / pml_size
        else:
            pos = (pml_size - i + 1) / pml_size
        val = math.exp(-(pml_alpha * pos ** 4) * dt)
        pml[i - 1] = val
        pml[N - i] = val
    return pml


def db2neper(alpha_coeff, alpha_power):
    """conversion.jl:34 -- alpha*(1e-2/(20 log10 e))*(2 pi 1e6)^-y (reference's 1e-2)."""
    return alpha_coeff * (1e-2 / (20 * math.log10(math.e))) * (2 * math.pi * 1e6) ** (-alpha_power)


def compute_kappa(k: np.ndarray, c_ref: float, dt: float) -> np.ndarray:
    """first_order.jl:35-42 -- sin(a)/a, a = c_ref*k*dt/2, 1 where a == 0."""
    arg = c_ref * k * dt / 2
    out = np.ones_like(arg)
    nz = arg != 0
    out[nz] = np.sin(arg[nz]) / arg[nz]
    return out


def is_lossless(medium) -> bool:
    """medium.jl:90."""
    return medium.alpha_coeff is None or medium.alpha_mode == "no_absorption"


def precompute_absorption(medium, k: np.ndarray, c_ref: float):
    """first_order.jl:13-30 -> (tau, eta, nabla1, nabla2, mode) or None."""
    if is_lossless(medium):
        return None
    y = medium.alpha_power
    if y is None:
        raise TypeError("alpha_power must be set when alpha_coeff is set")
    if np.ndim(medium.alpha_coeff) != 0:
        # first_order.jl:588-592: T(absorb_tau) of an array throws
        raise TypeError("array alpha_coeff is not supported (scalar conversion fails)")
    a_np = db2neper(float(medium.alpha_coeff), float(y))
    tau = -2 * a_np * c_ref ** (y - 1)
    eta = 2 * a_np * c_ref ** (y - 1) * math.tan(math.pi * y / 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        nabla1 = tau * k ** y
        nabla2 = eta * k ** (y - 1)
    return dict(tau=tau, eta=eta, nabla1=nabla1, nabla2=nabla2, mode=medium.alpha_mode)


def stagger_density(rho0: np.ndarray, dim: int) -> np.ndarray:
    """first_order_steps.jl:40-52 -- (rho0 + circshift(rho0, -e_dim))/2 (periodic)."""
    return (rho0 + np.roll(rho0, -1, axis=dim)) / 2


# ---------------------------------------------------------------------------
# Input builders (signal/generation.jl, geometry/shapes.jl, filter/filters.jl)
# ---------------------------------------------------------------------------

def julia_round(x: float) -> int:
    """Julia's round(Int, x): ties to even (same as Python's round)."""
    return int(round(x))


def tone_burst(sample_freq: float, signal_freq: float, num_cycles: int,
               envelope: str = "gaussian", signal_offset: int = 0) -> np.ndarray:
    """generation.jl:30-71."""
    tone_length = num_cycles / signal_freq
    dt = 1.0 / sample_freq
    N = julia_round(tone_length / dt)
    t = np.arange(N) * dt
    carrier = np.sin(2 * math.pi * signal_freq * t)
    if envelope == "gaussian":
        t_center = tone_length / 2
        sigma = tone_length / (2 * num_cycles)
        env = np.exp(-((t - t_center) / sigma) ** 2 / 2)
    elif envelope == "rectangular":
        env = np.ones(N)
    elif envelope == "cosine":
        env = np.array([0.5 * (1.0 - math.cos(2 * math.pi * i / (N - 1))) for i in range(N)])
    else:
        raise ValueError(f"Unknown envelope type: {envelope}")
    sig = carrier * env
    if signal_offset > 0:
        sig = np.concatenate([np.zeros(signal_offset), sig])
    return sig


def make_disc(Nx, Ny, cx, cy, radius) -> np.ndarray:
    """shapes.jl:22-38 (1-based centre)."""
    cx, cy = julia_round(cx), julia_round(cy)
    i = np.arange(1, Nx + 1)[:, None]
    j = np.arange(1, Ny + 1)[None, :]
    return (i - cx) ** 2 + (j - cy) ** 2 <= radius ** 2


def make_circle(Nx, Ny, cx, cy, radius) -> np.ndarray:
    """shapes.jl:56-107 -- full circle by the midpoint algorithm."""
    cx, cy, r = julia_round(cx), julia_round(cy), julia_round(radius)
    mask = np.zeros((Nx, Ny), dtype=bool)

    def setp(px, py):
        if 1 <= px <= Nx and 1 <= py <= Ny:
            mask[px - 1, py - 1] = True

    x, y, dd = r, 0, 1 - r
    while x >= y:
        for (a, b) in ((cx + x, cy + y), (cx - x, cy + y), (cx + x, cy - y), (cx - x, cy - y),
                       (cx + y, cy + x), (cx - y, cy + x), (cx + y, cy - x), (cx - y, cy - x)):
            setp(a, b)
        y += 1
        if dd <= 0:
            dd += 2 * y + 1
        else:
            x -= 1
            dd += 2 * (y - x) + 1
    return mask


def make_ball(Nx, Ny, Nz, cx, cy, cz, radius) -> np.ndarray:
    """shapes.jl:260-278."""
    cx, cy, cz = julia_round(cx), julia_round(cy), julia_round(cz)
    i = np.arange(1, Nx + 1)[:, None, None]
    j = np.arange(1, Ny + 1)[None, :, None]
    k = np.arange(1, Nz + 1)[None, None, :]
    return (i - cx) ** 2 + (j - cy) ** 2 + (k - cz) ** 2 <= radius ** 2


def make_bowl(Nx, Ny, Nz, center, radius, diameter, focus) -> np.ndarray:
    """shapes.jl:330-377 -- shell |r-c| in [R-0.5, R+0.5] inside the aperture cone."""
    cx, cy, cz = (julia_round(v) for v in center)
    fx, fy, fz = (julia_round(v) for v in focus)
    dvec = np.array([fx - cx, fy - cy, fz - cz], dtype=np.float64)
    nrm = float(np.linalg.norm(dvec))
    bowl = np.zeros((Nx, Ny, Nz), dtype=bool)
    if nrm == 0.0:
        return bowl
    dvec /= nrm
    half_angle = math.asin(min(max(diameter / (2 * radius), -1.0), 1.0))
    r2i, r2o = (radius - 0.5) ** 2, (radius + 0.5) ** 2
    i = np.arange(1, Nx + 1)[:, None, None] - cx
    j = np.arange(1, Ny + 1)[None, :, None] - cy
    k = np.arange(1, Nz + 1)[None, None, :] - cz
    d2 = i ** 2 + j ** 2 + k ** 2
    shell = (d2 >= r2i) & (d2 <= r2o)
    with np.errstate(invalid="ignore", divide="ignore"):
        cosang = (i * dvec[0] + j * dvec[1] + k * dvec[2]) / np.sqrt(d2)
    return shell & (cosang >= math.cos(half_angle))


def blackman_window(N: int) -> np.ndarray:
    """filters.jl:67-73."""
    if N == 1:
        return np.array([1.0])
    n = np.arange(N)
    return 0.42 - 0.5 * np.cos(2 * math.pi * n / (N - 1)) + 0.08 * np.cos(4 * math.pi * n / (N - 1))


def smooth(mat: np.ndarray, restore_max: bool = False) -> np.ndarray:
    """filters.jl:20-60 -- per-axis FFT, multiply by fftshift(blackman), IFFT."""
    res = np.asarray(mat, dtype=np.complex128).copy()
    for d in range(res.ndim):
        N = res.shape[d]
        win = np.fft.fftshift(blackman_window(N))
        shape = [1] * res.ndim
        shape[d] = N
        res = np.fft.fft(res, axis=d)
        res = res * win.reshape(shape)
        res = np.fft.ifft(res, axis=d)
    out = res.real.copy()
    if restore_max:
        omax = float(np.max(np.abs(mat)))
        smax = float(np.max(np.abs(out)))
        if smax > 0:
            out *= omax / smax
    return out


# ---------------------------------------------------------------------------
# FFT plumbing (fft_utils.jl)
# ---------------------------------------------------------------------------

def _axes(nd):
    # rfftn halves the LAST listed axis; the reference halves x (dim 1).
    return tuple(range(nd - 1, -1, -1))


def rfft_nd(f: np.ndarray) -> np.ndarray:
    """fft_utils.jl:59,136 -- unnormalised rfft over all dims, x halved."""
    return sfft.rfftn(f, axes=_axes(f.ndim), workers=_WORKERS)


def irfft_nd(S: np.ndarray, shape) -> np.ndarray:
    """fft_utils.jl:60,179 -- normalised irfft (1/N) back to the real grid."""
    nd = len(shape)
    s = tuple(shape[a] for a in _axes(nd))
    return sfft.irfftn(S, s=s, axes=_axes(nd), workers=_WORKERS)


def spectral_gradient(f: np.ndarray, k_vec: np.ndarray, shift: np.ndarray, dim: int) -> np.ndarray:
    """fft_utils.jl:126-182 -- irfft(i k_d shift_d rfft(f)); k, shift sliced to Nx//2+1 on dim 0."""
    S = rfft_nd(f)
    n1 = S.shape[0]
    nd = f.ndim
    if dim == 0:
        kv, sv = k_vec[:n1], shift[:n1]
    else:
        kv, sv = k_vec, shift
    shape = [1] * nd
    shape[dim] = kv.shape[0]
    S = 1j * kv.reshape(shape) * sv.reshape(shape) * S
    return irfft_nd(S, f.shape)


# ---------------------------------------------------------------------------
# Equation of state (first_order_steps.jl:214-290)
# ---------------------------------------------------------------------------

def compute_pressure(rho_total, rho_total_prev, kappa_r, c0, rho0, BonA, absorb, dt, shape):
    is_absorbing = absorb is not None and absorb["mode"] != "no_absorption"
    is_nonlinear = BonA is not None
    if np.ndim(c0) == 0:
        c_sq = float(c0) ** 2
        if is_nonlinear:
            tmp = rho_total * (1 + BonA / 2 * rho_total / rho0)
            S1 = rfft_nd(tmp)
        else:
            tmp = None
            S1 = rfft_nd(rho_total)
        if not is_absorbing:
            S1 = c_sq * kappa_r * S1                                         # :245
        elif absorb["mode"] == "no_dispersion" or rho_total_prev is None:
            S1 = c_sq * kappa_r * (1 + absorb["nabla1_r"]) * S1               # :250
        else:
            if is_nonlinear:                                                  # :255-259
                rho_eff_prev = rho_total_prev * (1 + BonA / 2 * rho_total_prev / rho0)
                tmp = (tmp - rho_eff_prev) / dt
            else:
                tmp = (rho_total - rho_total_prev) / dt                       # :261
            S2 = rfft_nd(tmp)
            S1 = c_sq * kappa_r * ((1 + absorb["nabla1_r"]) * S1 - absorb["nabla2_r"] * S2)
        return irfft_nd(S1, shape)
    # heterogeneous c0 (:273-289): never dispersion
    if is_nonlinear:
        tmp = c0 ** 2 * rho_total * (1 + BonA / 2 * rho_total / rho0)
    else:
        tmp = c0 ** 2 * rho_total
    S1 = rfft_nd(tmp)
    if not is_absorbing:
        S1 = kappa_r * S1
    else:
        S1 = kappa_r * (1 + absorb["nabla1_r"]) * S1
    return irfft_nd(S1, shape)


# ---------------------------------------------------------------------------
# The solver (first_order.jl:92-728, first_order_steps.jl:58-825)
# ---------------------------------------------------------------------------

def _as_param(x):
    if x is None:
        return None
    a = np.asarray(x, dtype=np.float64)
    return float(a) if a.ndim == 0 else a


def _findall(mask: np.ndarray) -> np.ndarray:
    """Column-major linear indices of true entries (Julia findall order)."""
    return np.flatnonzero(np.asarray(mask, dtype=bool).ravel(order="F"))


def _norm_tuple(v, nd, cast):
    if isinstance(v, (tuple, list)):
        return tuple(cast(x) for x in v)
    return tuple(cast(v) for _ in range(nd))


def kspace_first_order(kgrid, medium, source, sensor, *, pml_size=20, pml_alpha=2.0,
                       smooth_p0=True, data_cast=np.float64, progress_callback=None,
                       n_steps: Optional[int] = None, axisymmetric: bool = False,
                       initial_state: Optional[Dict[str, Any]] = None,
                       **_ignored) -> Dict[str, np.ndarray]:
    """Restatement of kspace_first_order 1D/2D/3D (first_order.jl:92-281, 287-499, 505-728).

    ``kgrid``/``medium``/``source``/``sensor`` are duck-typed objects with the
    reference's field names.  ``n_steps`` (oracle-only) truncates the loop for
    timing samples.  ``axisymmetric`` selects the kspace_first_order_as loop
    (see that function).  ``initial_state`` (oracle-only: dict with ``p``,
    ``u`` and ``rho`` lists per axis, optional ``rho_prev``) replaces the
    state before step 1, so that n_steps=1 is one time_step_{1,2,3}d! from
    that state.  Returns the sensor-data dict (Float64 arrays).
    """
    g = grid_from(kgrid)
    nd = g.ndim
    shape = g.N
    T = np.dtype(data_cast).type
    if g.Nt == 0 or g.dt == 0:
        raise ValueError("Time array not set. Call make_time!(kgrid, sound_speed) before running.")
    record = [str(r) for r in (sensor.record or [])]
    for f in record:                                                          # sensor.jl:52-58
        if f not in VALID_RECORD_FIELDS:
            raise ValueError(f"Invalid sensor record field: :{f}")
    tr_data = getattr(sensor, "time_reversal_boundary_data", None)
    if axisymmetric:
        if nd != 2:
            raise ValueError("kspace_first_order_as needs a 2-D grid (axial x, radial y)")
        tr_data = None                 # axisymmetric.jl has no time-reversal branch
    is_tr = tr_data is not None
    dt, Nt = g.dt, g.Nt

    psz = _norm_tuple(pml_size, nd, int)
    pal = _norm_tuple(pml_alpha, nd, float)
    pml = [get_pml(g.N[a], g.d[a], psz[a], pal[a], dt) for a in range(nd)]
    pml_sg = [get_pml(g.N[a], g.d[a], psz[a], pal[a], dt, staggered=True) for a in range(nd)]

    c0 = _as_param(medium.sound_speed)
    rho0 = _as_param(medium.density)
    BonA = _as_param(getattr(medium, "BonA", None))
    c_ref = c0 if np.ndim(c0) == 0 else float(np.mean(c0))                  # first_order.jl:361
    n1 = shape[0] // 2 + 1
    kappa_r = compute_kappa(g.k, c_ref, dt)[:n1].astype(T)                    # :364
    absorb = precompute_absorption(medium, g.k, c_ref)
    if absorb is not None:
        absorb["nabla1_r"] = absorb["nabla1"][:n1].astype(T)
        absorb["nabla2_r"] = absorb["nabla2"][:n1].astype(T)

    def bshape(a):
        s = [1] * nd
        s[a] = -1
        return s

    pmlb = [pml[a].reshape(bshape(a)) for a in range(nd)]
    pmlsgb = [pml_sg[a].reshape(bshape(a)) for a in range(nd)]
    inv_r = None
    if axisymmetric:                                                          # axisymmetric.jl:103-106
        r_vec = np.abs(make_spatial_vec(shape[1], g.d[1]))
        r_vec[r_vec == 0] = g.d[1] / 2
        inv_r = (1.0 / r_vec).reshape(1, -1)
    rho0_sg = None
    if np.ndim(rho0) != 0:
        rho0_sg = [stagger_density(rho0, a).astype(T) for a in range(nd)]

    # state kept in Fortran order so column-major index writes are views
    p = np.zeros(shape, dtype=T, order="F")
    u = [np.zeros(shape, dtype=T, order="F") for _ in range(nd)]
    rho = [np.zeros(shape, dtype=T, order="F") for _ in range(nd)]
    rho_total_prev = None
    if absorb is not None and absorb["mode"] != "no_dispersion":
        rho_total_prev = np.zeros(shape, dtype=T, order="F")

    # sensor preallocation (sensor.jl:94-136)
    mask = getattr(sensor, "mask", None)
    data: Dict[str, np.ndarray] = {}
    mask_idx = np.zeros(0, dtype=np.int64)
    if mask is not None:
        mask_arr = np.asarray(mask)
        if mask_arr.dtype == bool:
            mask_idx = _findall(mask_arr)
            n_sensor = mask_idx.size
        else:
            n_sensor = mask_arr.shape[1]          # Cartesian: preallocated, never filled (A-13)
        for f in record:
            if f in ("p", "ux", "uy", "uz"):
                data[f] = np.zeros((n_sensor, Nt))
            elif f in ("p_max", "u_max", "I_max"):
                data[f] = np.full(n_sensor, -np.inf)
            elif f == "p_min":
                data[f] = np.full(n_sensor, np.inf)
            else:
                data[f] = np.zeros(n_sensor)

    tr_idx = np.zeros(0, dtype=np.int64)
    if is_tr:
        tr_data = np.asarray(tr_data, dtype=np.float64)
        Nt = tr_data.shape[1]                                                 # :213, 430, 658
        if mask is not None and np.asarray(mask).dtype == bool:
            tr_idx = _findall(mask)

    # p0 (first_order_steps.jl:58-191)
    p0 = getattr(source, "p0", None)
    if not is_tr and p0 is not None:
        p0 = np.asarray(p0, dtype=np.float64).astype(T)
        do_smooth = smooth_p0 and all(n > 1 for n in shape) if nd > 1 else (smooth_p0 and shape[0] > 4)
        if do_smooth:
            p0 = smooth(p0, restore_max=True).astype(T)
        p[...] = p0
        c0_sq = c0 ** 2
        for a in range(nd):
            rho[a][...] = p0 / (nd * c0_sq) if nd > 1 else p0 / c0_sq
        for a in range(nd):
            dp = spectral_gradient(p0, g.k_vecs[a], g.shifts[a][0], a)
            r = rho0 if rho0_sg is None else stagger_density(rho0, a)
            u[a][...] = -dt / (2 * r) * dp

    if initial_state is not None:
        p[...] = np.asarray(initial_state["p"], dtype=np.float64).astype(T)
        for a in range(nd):
            u[a][...] = np.asarray(initial_state["u"][a], dtype=np.float64).astype(T)
            rho[a][...] = np.asarray(initial_state["rho"][a], dtype=np.float64).astype(T)
        if rho_total_prev is not None and initial_state.get("rho_prev") is not None:
            rho_total_prev[...] = np.asarray(initial_state["rho_prev"], dtype=np.float64).astype(T)

    # sources (first_order_steps.jl:541-715)
    p_mask = getattr(source, "p_mask", None)
    p_sig = getattr(source, "p", None)
    has_p_src = p_mask is not None and p_sig is not None
    if has_p_src:
        p_src_idx = _findall(p_mask)
        p_sig = np.asarray(p_sig, dtype=np.float64)
        if p_sig.ndim == 1:
            p_sig = p_sig[:, None]
    u_mask = getattr(source, "u_mask", None)
    u_sigs = [getattr(source, n, None) for n in ("ux", "uy", "uz")[:nd]]
    if u_mask is not None:
        u_src_idx = _findall(u_mask)
        u_sigs = [None if s is None else np.asarray(s, dtype=np.float64) for s in u_sigs]
        u_sigs = [s if (s is None or s.ndim == 2) else s[:, None] for s in u_sigs]
    p_mode = int(getattr(source, "p_mode", ADDITIVE))
    u_mode = int(getattr(source, "u_mode", ADDITIVE))

    def set_at(arr, idx, vals, add):
        v = arr.ravel(order="F") if arr.flags.f_contiguous else None
        if v is None:
            tmp = arr.ravel(order="F").copy()
            if add:
                np.add.at(tmp, idx, vals)
            else:
                tmp[idx] = vals
            arr[...] = tmp.reshape(arr.shape, order="F")
        else:
            if add:
                np.add.at(v, idx, vals)
            else:
                v[idx] = vals

    c0_flat = None if np.ndim(c0) == 0 else c0.ravel(order="F")
    n_run = Nt if n_steps is None else min(Nt, n_steps)

    for t in range(1, n_run + 1):
        if is_tr and tr_idx.size:                                             # :234-242, 450-460, 678-689
            tr_t = Nt - t + 1
            cl = c0 if c0_flat is None else c0_flat[tr_idx]
            val = tr_data[:, tr_t - 1] / (nd * cl ** 2)
            for a in range(nd):
                set_at(rho[a], tr_idx, val.astype(T), add=False)

        # --- time_step_{1,2,3}d! (first_order_steps.jl:296-535)
        dp = [spectral_gradient(p, g.k_vecs[a], g.shifts[a][0], a) for a in range(nd)]
        for a in range(nd):
            r = rho0 if rho0_sg is None else rho0_sg[a]
            u[a][...] = pmlsgb[a] * (pmlsgb[a] * u[a] - dt / r * dp[a])
        if u_mask is not None:
            for a in range(nd):
                s = u_sigs[a]
                if s is None:
                    continue
                tc = min(t, s.shape[1])
                set_at(u[a], u_src_idx, s[:, tc - 1].astype(T), add=(u_mode != DIRICHLET))
        du = [spectral_gradient(u[a], g.k_vecs[a], g.shifts[a][1], a) for a in range(nd)]
        if inv_r is not None:                                                 # axisymmetric.jl:179-187
            du[1] = du[1] + u[1] * inv_r
        for a in range(nd):
            rho[a][...] = pmlb[a] * (pmlb[a] * rho[a] - dt * rho0 * du[a])
        if has_p_src:
            tc = min(t, p_sig.shape[1])
            cl = c0 if c0_flat is None else c0_flat[p_src_idx]
            val = (p_sig[:, tc - 1] / (nd * cl ** 2)).astype(T)
            for a in range(nd):
                set_at(rho[a], p_src_idx, val, add=(p_mode != DIRICHLET))
        rho_total = rho[0].copy(order="F")
        for a in range(1, nd):
            rho_total = rho_total + rho[a]
        p[...] = compute_pressure(rho_total, rho_total_prev, kappa_r, c0, rho0, BonA,
                                  absorb, dt, shape)
        if rho_total_prev is not None:
            rho_total_prev[...] = rho_total

        # --- record_sensor_data! (first_order_steps.jl:728-794)
        if mask_idx.size:
            _record(data, record, p, u, mask_idx, t, Nt)
        if progress_callback is not None:
            progress_callback(t, Nt, p)

    # finalize (first_order_steps.jl:818-825)
    if "p_rms" in data:
        data["p_rms"] = np.sqrt(data["p_rms"] / Nt)
    if "u_rms" in data:
        data["u_rms"] = np.sqrt(data["u_rms"] / Nt)
    if is_tr:
        data["p_final"] = p.ravel(order="F").astype(np.float64).copy()
    apply_sensor_directivity(data, sensor, g)
    apply_frequency_response(data, sensor, g)
    data["_state"] = dict(p=p, u=u, rho=rho)   # oracle-only: final fields for operator tests
    return data


def kspace_first_order_as(kgrid, medium, source, sensor, **kwargs) -> Dict[str, np.ndarray]:
    """Restatement of kspace_first_order_as (axisymmetric.jl:26-227).

    The 2-D loop with x axial and y radial, r = |y_vec| (0 replaced by dy/2,
    axisymmetric.jl:103-106), and the radial density update
    rho_r <- pml_r (pml_r rho_r - dt rho0 (d ur/dr + ur / r)) (179-187); no
    time-reversal branch.  Same keywords as kspace_first_order.
    """
    return kspace_first_order(kgrid, medium, source, sensor, axisymmetric=True, **kwargs)


def pstd_elastic(kgrid, medium, source, sensor, *, pml_size=None, pml_alpha=2.0, data_cast=np.float64,
                 progress_callback=None, **_ignored) -> Dict[str, np.ndarray]:
    """Restatement of pstd_elastic_2d / pstd_elastic_3d (elastic.jl:133-368, 392-588).

    Stress-velocity PSTD without k-space correction: v_a <- Psg_a (Psg_a v_a + dt/rho_sg,a sum_b
    D_b+ s_ab); normal stresses s_aa <- P_a (P_a s_aa + dt (lam div v + 2 mu D_a- v_a)); shear
    s_ab <- P_a P_b (s_ab + dt mu (D_b- v_a + D_a- v_b)).  2-D with scalar Lame parameters adds the
    Kelvin-Voigt factors (1 + chi_p), (1 + chi_s) (elastic.jl:303-316); 3-D has none.  p = -tr(s)/ndim.
    """
    g = grid_from(kgrid)
    nd = g.ndim
    if nd not in (2, 3):
        raise ValueError("pstd_elastic needs a 2-D or 3-D grid")
    shape = g.N
    T = np.dtype(data_cast).type
    if g.Nt == 0 or g.dt == 0:
        raise ValueError("Time array not set. Call make_time!(kgrid, sound_speed) before running.")
    record = [str(r) for r in (sensor.record or [])]
    for f in record:
        if f not in VALID_RECORD_FIELDS:
            raise ValueError(f"Invalid sensor record field: :{f}")
    dt, Nt = g.dt, g.Nt
    if pml_size is None:
        pml_size = 20 if nd == 2 else 10                                      # elastic.jl:137, 396
    psz = _norm_tuple(pml_size, nd, int)
    pal = _norm_tuple(pml_alpha, nd, float)

    def bs(v, ax):
        s = [1] * nd
        s[ax] = -1
        return v.reshape(s)

    P = [bs(get_pml(g.N[ax], g.d[ax], psz[ax], pal[ax], dt), ax) for ax in range(nd)]
    Psg = [bs(get_pml(g.N[ax], g.d[ax], psz[ax], pal[ax], dt, staggered=True), ax) for ax in range(nd)]
    cp = _as_param(medium.sound_speed_compression)
    cs = _as_param(medium.sound_speed_shear)
    rho0 = _as_param(medium.density)
    lam = rho0 * (cp ** 2 - 2 * cs ** 2)                                      # elastic.jl:186-190
    mu = rho0 * cs ** 2
    scalar_lame = np.ndim(lam) == 0
    chi_p = chi_s = 0.0
    y = getattr(medium, "alpha_power", None)
    if nd == 2 and y is not None:                                             # elastic.jl:193-210
        if getattr(medium, "alpha_coeff_compression", None) is not None:
            c_ref = cp if np.ndim(cp) == 0 else float(np.mean(cp))
            chi_p = 2 * db2neper(float(medium.alpha_coeff_compression), y) * c_ref ** (y - 1)
        if getattr(medium, "alpha_coeff_shear", None) is not None:
            cs_ref = cs if np.ndim(cs) == 0 else float(np.mean(cs))
            chi_s = 2 * db2neper(float(medium.alpha_coeff_shear), y) * cs_ref ** (y - 1)
    if not scalar_lame:
        chi_p = chi_s = 0.0                      # the array branch carries no viscous terms
    rho_sg = [rho0 if np.ndim(rho0) == 0 else stagger_density(rho0, ax) for ax in range(nd)]
    Dp = lambda f, ax: spectral_gradient(f, g.k_vecs[ax], g.shifts[ax][0], ax)
    Dm = lambda f, ax: spectral_gradient(f, g.k_vecs[ax], g.shifts[ax][1], ax)

    v = [np.zeros(shape, dtype=T) for _ in range(nd)]
    sn = [np.zeros(shape, dtype=T) for _ in range(nd)]                        # sxx, syy, (szz)
    pairs = [(0, 1)] if nd == 2 else [(0, 1), (0, 2), (1, 2)]                 # sxy, (sxz, syz)
    ss = {pr: np.zeros(shape, dtype=T) for pr in pairs}
    p0 = getattr(source, "p0", None)
    if p0 is not None:
        for ax in range(nd):
            sn[ax][...] = -np.asarray(p0, dtype=np.float64)

    mask = getattr(sensor, "mask", None)
    data: Dict[str, np.ndarray] = {}
    mask_idx = np.zeros(0, dtype=np.int64)
    if mask is not None:
        m = np.asarray(mask)
        n_sensor = _findall(m).size if m.dtype == bool else m.shape[1]
        if m.dtype == bool:
            mask_idx = _findall(m)
        for f in record:
            if f in ("p", "ux", "uy", "uz"):
                data[f] = np.zeros((n_sensor, Nt))
            elif f in ("p_max", "u_max", "I_max"):
                data[f] = np.full(n_sensor, -np.inf)
            elif f == "p_min":
                data[f] = np.full(n_sensor, np.inf)
            else:
                data[f] = np.zeros(n_sensor)

    def inject(arr, idx, sig, t, dirichlet):
        s = np.asarray(sig, dtype=np.float64)
        s = s[:, None] if s.ndim == 1 else s
        col = min(t, s.shape[1]) - 1
        flat = arr.ravel(order="F").copy()
        if dirichlet:
            flat[idx] = s[:, col]
        else:
            np.add.at(flat, idx, s[:, col])
        arr[...] = flat.reshape(arr.shape, order="F")

    u_mask = getattr(source, "u_mask", None)
    s_mask = getattr(source, "s_mask", None)
    u_idx = _findall(u_mask) if u_mask is not None else None
    s_idx = _findall(s_mask) if s_mask is not None else None
    u_dir = int(getattr(source, "u_mode", ADDITIVE)) == DIRICHLET
    s_dir = int(getattr(source, "s_mode", ADDITIVE)) == DIRICHLET

    def s_of(a_, b_):
        if a_ == b_:
            return sn[a_]
        return ss[(min(a_, b_), max(a_, b_))]

    for t in range(1, Nt + 1):
        for ax in range(nd):                                                  # momentum
            div = sum(Dp(s_of(ax, b), b) for b in range(nd))
            v[ax][...] = Psg[ax] * (Psg[ax] * v[ax] + dt / rho_sg[ax] * div)
        if u_idx is not None:
            for ax, nm in enumerate(("ux", "uy", "uz")[:nd]):
                sig = getattr(source, nm, None)
                if sig is not None:
                    inject(v[ax], u_idx, sig, t, u_dir)
        dv = [[Dm(v[ax], b) for b in range(nd)] for ax in range(nd)]          # dv[a][b] = d v_a / d x_b
        div_v = sum(dv[ax][ax] for ax in range(nd))
        for ax in range(nd):
            inner = lam * div_v + 2 * mu * dv[ax][ax]
            sn[ax][...] = P[ax] * (P[ax] * sn[ax] + dt * (1 + chi_p) * inner)
        for (a_, b_) in pairs:
            ss[(a_, b_)][...] = P[a_] * P[b_] * (ss[(a_, b_)] + dt * (1 + chi_s) * mu * (dv[a_][b_] + dv[b_][a_]))
        if s_idx is not None:
            names = ("sxx", "syy", "szz")[:nd]
            for ax, nm in enumerate(names):
                sig = getattr(source, nm, None)
                if sig is not None:
                    inject(sn[ax], s_idx, sig, t, s_dir)
            for (a_, b_) in pairs:
                nm = "s" + "xyz"[a_] + "xyz"[b_]
                sig = getattr(source, nm, None)
                if sig is not None:
                    inject(ss[(a_, b_)], s_idx, sig, t, s_dir)
        p = -sum(sn) / nd
        if mask_idx.size:
            _record(data, record, p, v, mask_idx, t, Nt)
        if progress_callback is not None:
            progress_callback(t, Nt, p)
    if "p_rms" in data:
        data["p_rms"] = np.sqrt(data["p_rms"] / Nt)
    if "u_rms" in data:
        data["u_rms"] = np.sqrt(data["u_rms"] / Nt)
    return data


def _record(data, record, p, u, idx, t, Nt):
    pf = p.ravel(order="F")
    uf = [a.ravel(order="F") for a in u]
    pv = pf[idx].astype(np.float64)
    for f in record:
        if f == "p":
            data["p"][:, t - 1] = pv
        elif f == "p_max":
            data["p_max"] = np.maximum(data["p_max"], pv)
        elif f == "p_min":
            data["p_min"] = np.minimum(data["p_min"], pv)
        elif f == "p_rms":
            data["p_rms"] += (pf[idx] ** 2).astype(np.float64)
        elif f == "p_final" and t == Nt:
            data["p_final"][:] = pv
        elif f in ("ux", "uy", "uz"):
            a = "xyz".index(f[1])
            if a < len(uf):
                data[f][:, t - 1] = uf[a][idx]
        elif f in ("u_max", "u_rms", "u_final"):
            mag = np.sqrt(sum((a[idx] ** 2).astype(np.float64) for a in uf))
            if f == "u_max":
                data[f] = np.maximum(data[f], mag)
            elif f == "u_rms":
                data[f] += mag ** 2
            elif t == Nt:
                data[f][:] = mag
        elif f in ("I_avg", "I_max"):
            if p.dtype != np.float64:
                # first_order_steps.jl:804: _acoustic_intensity(p_val::Float64, ...)
                raise TypeError("I_avg/I_max need Float64 data (MethodError in the reference)")
            inst = np.sqrt(sum((pv * a[idx]) ** 2 for a in uf))
            if f == "I_avg":
                data[f] += inst / Nt
            else:
                data[f] = np.maximum(data[f], inst)


def _fftfreq(N, Fs):
    """filters.jl:158-165."""
    if N % 2 == 0:
        f = np.concatenate([np.arange(0, N // 2), np.arange(-(N // 2), 0)])
    else:
        f = np.concatenate([np.arange(0, (N - 1) // 2 + 1), np.arange(-((N - 1) // 2), 0)])
    return f * (Fs / N)


def apply_sensor_directivity(data, sensor, g: Grid):
    """first_order_steps.jl:836-873."""
    ang = getattr(sensor, "directivity_angle", None)
    size = getattr(sensor, "directivity_size", None)
    if ang is None or size is None or "p" not in data:
        return
    P = data["p"]
    n_sensor, Nt = P.shape
    Fs = 1.0 / g.dt
    freqs = _fftfreq(Nt, Fs)
    dmin = g.d[0] if g.ndim == 1 else min(g.d)
    ang = np.asarray(ang, dtype=np.float64)
    for j in range(n_sensor):
        a = float(ang) if ang.ndim == 0 else float(ang.ravel()[j])
        if math.isclose(a, 0.0, rel_tol=math.sqrt(np.finfo(float).eps), abs_tol=0.0):
            continue
        F = np.fft.fft(P[j, :])
        kv = 2 * math.pi * np.abs(freqs) / dmin
        arg = kv * size * math.sin(a) / 2
        dirc = np.ones_like(arg)
        nz = arg != 0
        dirc[nz] = np.sin(arg[nz]) / arg[nz]
        P[j, :] = np.fft.ifft(F * dirc).real


def apply_frequency_response(data, sensor, g: Grid):
    """first_order_steps.jl:880-910."""
    fr = getattr(sensor, "frequency_response", None)
    if fr is None or "p" not in data:
        return
    fc, bw = fr
    P = data["p"]
    n_sensor, Nt = P.shape
    freqs = _fftfreq(Nt, 1.0 / g.dt)
    resp = np.exp(-((np.abs(freqs) - fc) ** 2) / (2 * bw ** 2))
    for j in range(n_sensor):
        P[j, :] = np.fft.ifft(np.fft.fft(P[j, :]) * resp).real


# ---------------------------------------------------------------------------
# Discrete adjoint of the time loop w.r.t. the sound-speed map (no reference
# counterpart: KWave.jl has no gradient).  Scope: heterogeneous c0 map, linear
# EOS, lossless or absorbing without dispersion (the heterogeneous branch,
# first_order_steps.jl:273-289), additive pressure sources, p0 allowed,
# PML and heterogeneous rho0 allowed.  Loss L = 1/2 sum_{n,j} (p^n_j - obs_{j,n})^2
# over the recorded sensor points (n = 1..Nt).
#
# With D+^T = -D- (the staggered multipliers are conjugates of each other),
# K = F^-1 kappa(1+nabla1) F symmetric, and per step n (state n-1 -> n)
#   u_d^n   = Psg_d (Psg_d u_d^{n-1} - dt/rho0sg_d D+_d p^{n-1})
#   rho_d^n = P_d (P_d rho_d^{n-1} - dt rho0 D-_d u_d^n) + s_n / (ndim c^2) [src]
#   p^n     = K (c^2 sum_d rho_d^n)
# the reverse sweep is
#   Pbar   += M^T (p^n - obs_n)
#   Q = K Pbar, Q' = (dK/dc_ref) Pbar ; g += 2 c rho_sum^n Q ; gref += <c^2 rho_sum^n, Q'>
#   Rtot_d = Rbar_d + c^2 Q ; g[src] -= 2 s_n / (ndim c^3) sum_d Rtot_d[src]
#   Utot_d = Ubar_d + D+_d (dt rho0 P_d Rtot_d) ; Rbar_d <- P_d^2 Rtot_d
#   Pbar   = sum_d D-_d (dt/rho0sg_d Psg_d Utot_d) ; Ubar_d <- Psg_d^2 Utot_d
# and finally g += gref / N (c_ref = mean(c0)) and, for p0 runs,
# g -= 2 p0 / (ndim c^3) sum_d Rbar_d^0.
# ---------------------------------------------------------------------------

def _eos_multiplier(g: Grid, medium, c_ref, dt, derivative=False):
    """kappa(1+nabla1) on the rfft grid, or its derivative w.r.t. c_ref."""
    k = g.k
    a = c_ref * k * dt / 2
    with np.errstate(invalid="ignore", divide="ignore"):
        kap = np.where(a == 0, 1.0, np.sin(a) / np.where(a == 0, 1.0, a))
        dkap = np.where(a == 0, 0.0, (a * np.cos(a) - np.sin(a)) / np.where(a == 0, 1.0, a) ** 2 * (k * dt / 2))
    absorb = precompute_absorption(medium, k, c_ref)
    n1 = np.zeros_like(k)
    dn1 = np.zeros_like(k)
    if absorb is not None:
        y = medium.alpha_power
        a_np = db2neper(float(medium.alpha_coeff), y)
        n1 = absorb["nabla1"]
        with np.errstate(invalid="ignore", divide="ignore"):
            dn1 = -2 * a_np * (y - 1) * c_ref ** (y - 2) * k ** y
    m = dkap * (1 + n1) + kap * dn1 if derivative else kap * (1 + n1)
    return m[: g.N[0] // 2 + 1]


def grad_c0(kgrid, medium, source, sensor, p_obs, *, pml_size=20, pml_alpha=2.0, smooth_p0=False):
    """Loss and dL/dc0 by the discrete adjoint (FP64; small grids)."""
    g = grid_from(kgrid)
    nd, shape, dt, Nt = g.ndim, g.N, g.dt, g.Nt
    c0 = np.asarray(medium.sound_speed, dtype=np.float64)
    if c0.ndim == 0:
        raise ValueError("grad_c0 needs a sound-speed map")
    if getattr(medium, "BonA", None) is not None:
        raise NotImplementedError("nonlinear medium")
    if medium.alpha_coeff is not None and medium.alpha_mode not in ("no_absorption",):
        pass  # heterogeneous c0 never uses dispersion; absorption handled by the multiplier
    if getattr(source, "p_mode", ADDITIVE) == DIRICHLET and getattr(source, "p", None) is not None:
        raise NotImplementedError("Dirichlet sources")
    rho0 = _as_param(medium.density)
    c_ref = float(np.mean(c0))
    psz = _norm_tuple(pml_size, nd, int)
    pal = _norm_tuple(pml_alpha, nd, float)

    def bs(v, a):
        s = [1] * nd
        s[a] = -1
        return v.reshape(s)

    P = [bs(get_pml(g.N[a], g.d[a], psz[a], pal[a], dt), a) for a in range(nd)]
    Psg = [bs(get_pml(g.N[a], g.d[a], psz[a], pal[a], dt, staggered=True), a) for a in range(nd)]
    r0sg = [rho0 if np.ndim(rho0) == 0 else stagger_density(rho0, a) for a in range(nd)]
    M1 = _eos_multiplier(g, medium, c_ref, dt)
    M2 = _eos_multiplier(g, medium, c_ref, dt, derivative=True)
    K = lambda f, m: irfft_nd(m * rfft_nd(f), shape)
    Dp = lambda f, a: spectral_gradient(f, g.k_vecs[a], g.shifts[a][0], a)
    Dm = lambda f, a: spectral_gradient(f, g.k_vecs[a], g.shifts[a][1], a)
    c2 = c0 ** 2
    flat = lambda x: x.ravel(order="F")
    sens = _findall(sensor.mask)
    src_idx = _findall(source.p_mask) if getattr(source, "p_mask", None) is not None else np.zeros(0, np.int64)
    sig = np.asarray(source.p, dtype=np.float64) if src_idx.size else np.zeros((0, 1))
    if sig.ndim == 1:
        sig = sig[:, None]

    # forward, keeping rho_sum^n (small grids)
    p = np.zeros(shape)
    u = [np.zeros(shape) for _ in range(nd)]
    rho = [np.zeros(shape) for _ in range(nd)]
    p0 = getattr(source, "p0", None)
    if p0 is not None:
        p0 = np.asarray(p0, dtype=np.float64)
        if smooth_p0:
            p0 = smooth(p0, restore_max=True)
        p = p0.copy()
        rho = [p0 / (nd * c2) for _ in range(nd)]
        u = [-dt / (2 * r0sg[a]) * Dp(p0, a) for a in range(nd)]
    rsum, y = [], np.zeros((sens.size, Nt))
    for n in range(1, Nt + 1):
        for a in range(nd):
            u[a] = Psg[a] * (Psg[a] * u[a] - dt / r0sg[a] * Dp(p, a))
        for a in range(nd):
            rho[a] = P[a] * (P[a] * rho[a] - dt * rho0 * Dm(u[a], a))
        if src_idx.size:
            col = min(n, sig.shape[1]) - 1
            val = sig[:, col] / (nd * flat(c2)[src_idx])
            for a in range(nd):
                fa = flat(rho[a]).copy()
                np.add.at(fa, src_idx, val)
                rho[a] = fa.reshape(shape, order="F")
        rs = sum(rho)
        rsum.append(rs)
        p = K(c2 * rs, M1)
        y[:, n - 1] = flat(p)[sens]
    res = y - np.asarray(p_obs, dtype=np.float64)
    loss = 0.5 * float(np.sum(res ** 2))

    grad = np.zeros(shape)
    gref = 0.0
    Pb = np.zeros(shape)
    Ub = [np.zeros(shape) for _ in range(nd)]
    Rb = [np.zeros(shape) for _ in range(nd)]
    for n in range(Nt, 0, -1):
        fb = flat(Pb).copy()
        np.add.at(fb, sens, res[:, n - 1])
        Pb = fb.reshape(shape, order="F")
        Q = K(Pb, M1)
        Qd = K(Pb, M2)
        grad += 2 * c0 * rsum[n - 1] * Q
        gref += float(np.sum(c2 * rsum[n - 1] * Qd))
        Rt = [Rb[a] + c2 * Q for a in range(nd)]
        if src_idx.size:
            col = min(n, sig.shape[1]) - 1
            tot = sum(flat(Rt[a])[src_idx] for a in range(nd))
            gf = flat(grad).copy()
            np.add.at(gf, src_idx, -2 * sig[:, col] / (nd * flat(c0)[src_idx] ** 3) * tot)
            grad = gf.reshape(shape, order="F")
        Ut = [Ub[a] + Dp(dt * rho0 * P[a] * Rt[a], a) for a in range(nd)]
        Rb = [P[a] * P[a] * Rt[a] for a in range(nd)]
        Pb = sum(Dm(dt / r0sg[a] * Psg[a] * Ut[a], a) for a in range(nd))
        Ub = [Psg[a] * Psg[a] * Ut[a] for a in range(nd)]
    if p0 is not None:
        grad += -2 * p0 / (nd * c0 ** 3) * sum(Rb)
    grad += gref / c0.size
    return loss, grad
