"""Pins of the oracle's elastic solvers (pstd_elastic_2d/3d, elastic.jl:133-588).

- per Fourier mode, a y-invariant compressional (vx, sxx, syy) or shear
  (vy, sxy) plane wave evolves by an exact 2x2 leapfrog recurrence
  (v' = v + dt/rho D+ s; s' = s + dt M D- v') computed here by matrix powers;
  planted mistakes (D- in the momentum equation, lambda in place of
  lambda + 2 mu, the old velocity in the stress update) must fail;
- one step on a 6x5 heterogeneous grid against dense DFT matrices;
- the reference's own assertions (test_elastic.jl).
"""
import copy
import math

import numpy as np
import pytest

import kwave_amd as K
from oracle import kwave_oracle as O


def _plane(kind, N=32, nt=40, chi=False):
    dx = 1e-4
    kg = K.KWaveGrid(N, dx, 8, dx)
    cp, cs, rho = 3000.0, 1400.0, 1800.0
    K.make_time(kg, cp)
    kg.Nt = nt
    kg.t_array = np.arange(nt) * kg.dt
    kw = dict(sound_speed_compression=cp, sound_speed_shear=cs, density=rho)
    if chi:
        kw.update(alpha_coeff_compression=1e10, alpha_coeff_shear=2e10, alpha_power=1.5)   # chi_s ~ 0.1
    med = K.ElasticMedium(**kw)
    m = 3                                              # mode index along x
    x = np.arange(N)
    prof = np.cos(2 * np.pi * m * x / N)[:, None] * np.ones((1, 8))
    src = K.ElasticSource(p0=prof) if kind == "p" else K.ElasticSource(
        s_mask=np.ones((N, 8), bool), sxy=prof.ravel(order="F")[:, None] * np.r_[1.0, np.zeros(nt - 1)][None, :])
    sen = K.KWaveSensor(mask=np.ones((N, 8), bool), record=["p", "ux", "uy"])
    out = O.pstd_elastic(copy.deepcopy(kg), med, src, sen, pml_size=2, pml_alpha=0.0)
    return kg, med, out, m


def _mode(traces, N, m):
    """Complex amplitude of mode m along x (y-invariant data, column 0 of y)."""
    f = traces.reshape(N, 8, -1, order="F")[:, 0, :]
    return np.fft.fft(f, axis=0)[m] / N * 2


def test_compressional_plane_wave_recurrence():
    N = 32
    kg, med, out, m = _plane("p", N)
    dx, dt, nt = 1e-4, kg.dt, kg.Nt
    k = 2 * np.pi * m / (N * dx)
    Dp, Dm = 1j * k * np.exp(1j * k * dx / 2), 1j * k * np.exp(-1j * k * dx / 2)
    lam, mu = 1800.0 * (3000.0 ** 2 - 2 * 1400.0 ** 2), 1800.0 * 1400.0 ** 2
    v, sxx, syy = 0j, -1.0 + 0j, -1.0 + 0j                   # p0 = cos: sxx = syy = -p0
    vx_ref, p_ref = [], []
    for _ in range(nt):
        v = v + dt / 1800.0 * Dp * sxx
        sxx = sxx + dt * (lam + 2 * mu) * Dm * v
        syy = syy + dt * lam * Dm * v
        vx_ref.append(v)
        p_ref.append(-(sxx + syy) / 2)
    vx = _mode(out["ux"], N, m)
    p = _mode(out["p"], N, m)
    scale = np.abs(np.asarray(vx_ref)).max()
    assert np.abs(vx - np.asarray(vx_ref)).max() <= 1e-10 * scale
    assert np.abs(p - np.asarray(p_ref)).max() <= 1e-10 * np.abs(np.asarray(p_ref)).max()
    # planted mistakes: momentum with D-, lambda in place of lambda + 2 mu, old velocity
    for bad in ("dm", "lam", "old"):
        v, sxx, syy, w = 0j, -1.0 + 0j, -1.0 + 0j, []
        for _ in range(nt):
            vn = v + dt / 1800.0 * (Dm if bad == "dm" else Dp) * sxx
            vs = v if bad == "old" else vn
            sxx = sxx + dt * (lam if bad == "lam" else lam + 2 * mu) * Dm * vs
            v = vn
            w.append(v)
        assert np.abs(vx - np.asarray(w)).max() > 1e-4 * scale, bad


@pytest.mark.parametrize("chi", [False, True], ids=["elastic", "kelvin_voigt"])
def test_shear_plane_wave_recurrence(chi):
    N = 32
    kg, med, out, m = _plane("s", N, chi=chi)
    dx, dt, nt = 1e-4, kg.dt, kg.Nt
    k = 2 * np.pi * m / (N * dx)
    Dp, Dm = 1j * k * np.exp(1j * k * dx / 2), 1j * k * np.exp(-1j * k * dx / 2)
    mu = 1800.0 * 1400.0 ** 2
    chi_s = 2 * O.db2neper(2e10, 1.5) * 1400.0 ** 0.5 if chi else 0.0
    assert not chi or 0.05 < chi_s < 0.5
    v, sxy, vy_ref = 0j, 0j, []
    for t in range(nt):
        v = v + dt / 1800.0 * Dp * sxy
        sxy = sxy + dt * (1 + chi_s) * mu * Dm * v
        if t == 0:
            sxy = sxy + 1.0                                   # additive stress source, first step only
        vy_ref.append(v)
    vy = _mode(out["uy"], N, m)
    ref = np.asarray(vy_ref)
    assert np.abs(vy - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.abs(_mode(out["ux"], N, m)).max() <= 1e-12 * np.abs(ref).max()   # no compressional part


def test_one_step_dense_brute_force_2d():
    shape, hs = (6, 5), (1e-4, 1.2e-4)
    rng = np.random.default_rng(9)
    kg = K.KWaveGrid(6, 1e-4, 5, 1.2e-4)
    cp = 2500 + 500 * rng.random(shape)
    cs = 1000 + 300 * rng.random(shape)
    rho = 1500 + 300 * rng.random(shape)
    K.make_time(kg, cp)
    kg.Nt = 1
    med = K.ElasticMedium(sound_speed_compression=cp, sound_speed_shear=cs, density=rho)
    p0 = rng.standard_normal(shape)
    src = K.ElasticSource(p0=p0)
    sen = K.KWaveSensor(mask=np.ones(shape, bool), record=["p", "ux", "uy"])
    out = O.pstd_elastic(copy.deepcopy(kg), med, src, sen, pml_size=2, pml_alpha=3e6)
    dt = kg.dt
    ks = [O.make_wavenumber_vec(6, 1e-4), O.make_wavenumber_vec(5, 1.2e-4)]

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

    bs = lambda v, a: v.reshape([-1 if i == a else 1 for i in range(2)])
    P = [bs(O.get_pml(n, h, 2, 3e6, dt), a) for a, (n, h) in enumerate(zip(shape, hs))]
    Ps = [bs(O.get_pml(n, h, 2, 3e6, dt, staggered=True), a) for a, (n, h) in enumerate(zip(shape, hs))]
    lam, mu = rho * (cp ** 2 - 2 * cs ** 2), rho * cs ** 2
    sxx, syy, sxy = -p0, -p0, np.zeros(shape)
    vx = Ps[0] * (dt / O.stagger_density(rho, 0) * (deriv(0, 1)(sxx) + deriv(1, 1)(sxy)))
    vy = Ps[1] * (dt / O.stagger_density(rho, 1) * (deriv(0, 1)(sxy) + deriv(1, 1)(syy)))
    dxx, dyy = deriv(0, -1)(vx), deriv(1, -1)(vy)
    sxx = P[0] * (P[0] * sxx + dt * ((lam + 2 * mu) * dxx + lam * dyy))
    syy = P[1] * (P[1] * syy + dt * (lam * dxx + (lam + 2 * mu) * dyy))
    p = -(sxx + syy) / 2
    f = lambda a: a.ravel(order="F")
    assert np.allclose(out["ux"][:, 0], f(vx), rtol=1e-11, atol=1e-12 * np.abs(vx).max())
    assert np.allclose(out["uy"][:, 0], f(vy), rtol=1e-11, atol=1e-12 * np.abs(vy).max())
    assert np.allclose(out["p"][:, 0], f(p), rtol=1e-11, atol=1e-12 * np.abs(p).max())


def test_reference_assertions():
    em = K.ElasticMedium(sound_speed_compression=3000.0, sound_speed_shear=1500.0, density=1800.0)
    assert em.sound_speed_compression == 3000.0 and em.alpha_coeff_compression is None
    es = K.ElasticSource()
    assert es.s_mask is None and es.p0 is None and es.s_mode == K.Additive
    kg = K.KWaveGrid(32, 0.5e-3, 32, 0.5e-3)
    K.make_time(kg, 3000.0)
    p0 = np.zeros((32, 32))
    p0[15, 15] = 1.0
    mask = np.zeros((32, 32), bool)
    mask[4, :] = True
    r = O.pstd_elastic(copy.deepcopy(kg), em, K.ElasticSource(p0=p0), K.KWaveSensor(mask=mask, record=["p"]))
    assert r["p"].shape == (32, kg.Nt)
    sm = np.zeros((32, 32), bool)
    sm[4, 15] = True
    sig = np.sin(2 * np.pi * 1e6 * np.arange(kg.Nt) * kg.dt)[None, :]
    mask2 = np.zeros((32, 32), bool)
    mask2[15, :] = True
    r = O.pstd_elastic(copy.deepcopy(kg), em, K.ElasticSource(s_mask=sm, sxx=sig), K.KWaveSensor(mask=mask2, record=["p"]))
    assert not np.all(r["p"] == 0)
    kg3 = K.KWaveGrid(16, 0.5e-3, 16, 0.5e-3, 16, 0.5e-3)
    K.make_time(kg3, 3000.0)
    p0 = np.zeros((16, 16, 16))
    p0[7, 7, 7] = 1.0
    m3 = np.zeros((16, 16, 16), bool)
    m3[2, :, :] = True
    r = O.pstd_elastic(copy.deepcopy(kg3), em, K.ElasticSource(p0=p0), K.KWaveSensor(mask=m3, record=["p"]), pml_size=5)
    assert r["p"].shape[0] == 256
