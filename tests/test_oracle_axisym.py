"""Pins of the oracle's axisymmetric loop (kspace_first_order_as, axisymmetric.jl:26-227).

- one step on a 6x5 grid against dense matrices built from explicit DFT sums,
  with r = |y_vec| and the r = 0 point replaced by dy/2 (odd Ny puts a point
  on the axis); variants with a wrong radius (signed y, no axis fix, no 1/r
  term) must not match;
- a field that does not vary along y has no radial velocity, so the
  axisymmetric run equals the plain 2-D run;
- the reference's own assertions (test_axisymmetric.jl:2-55).
"""
import copy
import math

import numpy as np
import pytest

import kwave_amd as K
from oracle import kwave_oracle as O


def _dense_step(shape, hs, kg, c0, rho0, p0, r_kind):
    dt = kg.dt
    ks = [O.make_wavenumber_vec(shape[0], hs[0]), O.make_wavenumber_vec(shape[1], hs[1])]

    def dft(N):
        n = np.arange(N)
        return np.exp(-2j * math.pi * np.outer(n, n) / N)

    def deriv(a, sign):
        N = shape[a]
        D = 1j * ks[a] * np.exp(sign * 1j * ks[a] * hs[a] / 2)
        if N % 2 == 0:
            D[N // 2] = D[N // 2].real
        F = dft(N)
        M = (np.conj(F) / N @ np.diag(D) @ F).real
        return lambda f: np.moveaxis(np.tensordot(M, np.moveaxis(f, a, 0), axes=1), 0, a)

    pml = [O.get_pml(n, h, 2, 3e6, dt) for n, h in zip(shape, hs)]
    pmls = [O.get_pml(n, h, 2, 3e6, dt, staggered=True) for n, h in zip(shape, hs)]
    bs = lambda v, a: v.reshape([-1 if i == a else 1 for i in range(2)])
    y = (np.arange(shape[1]) - (shape[1] - 1) / 2) * hs[1]
    if r_kind == "abs_fix":
        r = np.abs(y)
        r[r == 0] = hs[1] / 2
        radial = lambda ur: ur / r[None, :]
    elif r_kind == "signed":
        r = y.copy()
        r[r == 0] = hs[1] / 2
        radial = lambda ur: ur / r[None, :]
    elif r_kind == "no_fix":
        r = np.abs(y)
        r[r == 0] = hs[1]
        radial = lambda ur: ur / r[None, :]
    else:
        radial = lambda ur: 0 * ur
    u = [-dt / (2 * O.stagger_density(rho0, a)) * deriv(a, +1)(p0) for a in range(2)]
    rho = [p0 / (2 * c0 ** 2) for _ in range(2)]
    for a in range(2):
        u[a] = bs(pmls[a], a) * (bs(pmls[a], a) * u[a] - dt / O.stagger_density(rho0, a) * deriv(a, +1)(p0))
    du = [deriv(a, -1)(u[a]) for a in range(2)]
    du[1] = du[1] + radial(u[1])
    for a in range(2):
        rho[a] = bs(pml[a], a) * (bs(pml[a], a) * rho[a] - dt * rho0 * du[a])
    f = c0 ** 2 * (rho[0] + rho[1])
    kk = np.sqrt(ks[0][:, None] ** 2 + ks[1][None, :] ** 2)
    c_ref = float(np.mean(c0))
    x = c_ref * kk * dt / 2
    kap = np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1, x))
    F0, F1 = dft(shape[0]), dft(shape[1])
    S = F0 @ f @ F1.T
    return (np.conj(F0) / shape[0] @ (kap * S) @ np.conj(F1).T / shape[1]).real


def test_one_step_dense_brute_force_axisymmetric():
    shape, hs = (6, 5), (1e-4, 1.2e-4)
    rng = np.random.default_rng(4)
    kg = K.KWaveGrid(6, 1e-4, 5, 1.2e-4)
    c0 = 1500 + 100 * rng.random(shape)
    rho0 = 1000 + 100 * rng.random(shape)
    K.make_time(kg, c0)
    kg.Nt = 1
    med = K.KWaveMedium(sound_speed=c0, density=rho0)
    p0 = rng.standard_normal(shape) * 1e5
    src = K.KWaveSource(p0=p0)
    sen = K.KWaveSensor(mask=np.ones(shape, bool), record=["p"])
    out = O.kspace_first_order_as(copy.deepcopy(kg), med, src, sen, pml_size=2, pml_alpha=3e6, smooth_p0=False)
    got = out["p"][:, 0]
    ref = _dense_step(shape, hs, kg, c0, rho0, p0, "abs_fix").ravel(order="F")
    tol = 1e-9 * np.abs(ref).max()
    assert np.allclose(got, ref, rtol=1e-11, atol=tol)
    for wrong in ("signed", "no_fix", "none"):
        bad = _dense_step(shape, hs, kg, c0, rho0, p0, wrong).ravel(order="F")
        assert np.abs(got - bad).max() > 1e3 * tol, wrong


def test_axial_plane_wave_equals_plain_2d():
    kg = K.KWaveGrid(64, 1e-4, 32, 1e-4)
    K.make_time(kg, 1500.0)
    kg.Nt = 60
    kg.t_array = np.arange(60) * kg.dt
    med = K.KWaveMedium(sound_speed=1500.0, density=1000.0, alpha_coeff=0.5, alpha_power=1.5,
                        alpha_mode="no_dispersion")
    x = np.arange(64)
    p0 = np.repeat(np.exp(-((x - 32) / 4.0) ** 2)[:, None], 32, axis=1)
    sen = K.KWaveSensor(mask=np.ones((64, 32), bool), record=["p", "p_final"])
    kw = dict(pml_size=(8, 0), pml_alpha=2.0, smooth_p0=False)
    a = O.kspace_first_order_as(copy.deepcopy(kg), med, K.KWaveSource(p0=p0), sen, **kw)
    b = O.kspace_first_order(copy.deepcopy(kg), med, K.KWaveSource(p0=p0), sen, **kw)
    assert np.abs(a["p"] - b["p"]).max() <= 1e-12 * np.abs(b["p"]).max()
    # a radially varying field does feel the 1/r term
    p1 = np.exp(-((x[:, None] - 32) ** 2 + (np.arange(32)[None, :] - 16) ** 2) / 16.0)
    a = O.kspace_first_order_as(copy.deepcopy(kg), med, K.KWaveSource(p0=p1), sen, **kw)
    b = O.kspace_first_order(copy.deepcopy(kg), med, K.KWaveSource(p0=p1), sen, **kw)
    assert np.abs(a["p"] - b["p"]).max() > 1e-2 * np.abs(b["p"]).max()


def test_reference_assertions():
    """test_axisymmetric.jl:2-55: outputs exist with the reference's shapes."""
    Nx, Nr, dx = 64, 32, 1e-4
    kg = K.KWaveGrid(Nx, dx, Nr, dx)
    med = K.KWaveMedium(sound_speed=1500.0, density=1000.0)
    K.make_time(kg, 1500.0, t_end=2e-6)
    i, j = np.meshgrid(np.arange(1, Nx + 1), np.arange(1, Nr + 1), indexing="ij")
    p0 = np.exp(-(((i - Nx // 2) * dx) ** 2 + ((j - Nr // 2) * dx) ** 2) / (5 * dx) ** 2)
    mask = np.zeros((Nx, Nr), bool)
    mask[Nx // 2 - 1, :] = True
    res = O.kspace_first_order_as(copy.deepcopy(kg), med, K.KWaveSource(p0=p0), K.KWaveSensor(mask=mask, record=["p"]))
    assert "p" in res and res["p"].shape[1] == kg.Nt and np.all(np.isfinite(res["p"]))
    kg2 = K.KWaveGrid(64, dx, 64, dx)
    med2 = K.KWaveMedium(sound_speed=1500.0, density=1000.0, alpha_coeff=0.5, alpha_power=1.5,
                         alpha_mode="no_dispersion")
    K.make_time(kg2, 1500.0, t_end=1e-6)
    p0 = np.zeros((64, 64))
    p0[31, 31] = 1.0
    m2 = np.zeros((64, 64), bool)
    m2[31, 0] = True
    res = O.kspace_first_order_as(copy.deepcopy(kg2), med2, K.KWaveSource(p0=p0),
                                  K.KWaveSensor(mask=m2, record=["p_max"]), pml_size=10)
    assert "p_max" in res


def test_needs_2d_grid():
    kg = K.KWaveGrid(16, 1e-4, 16, 1e-4, 16, 1e-4)
    K.make_time(kg, 1500.0)
    with pytest.raises(ValueError):
        O.kspace_first_order_as(kg, K.KWaveMedium(sound_speed=1500.0, density=1000.0),
                                K.KWaveSource(p0=np.zeros((16, 16, 16))), K.KWaveSensor(mask=None, record=["p"]))
