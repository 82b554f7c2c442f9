"""Stacked firings (kw_sim_create_firings) against one simulation per firing, through the C ABI.

A stack runs every firing's passes in the same launches; each firing is an
independent 2-D problem on the same grid and medium.  Each firing's sensor
data must equal its own kspace_first_order run (the same per-line arithmetic,
so FP64 agrees to 1e-13 and FP32 to 1e-6 relative L2), and the stacked
adjoint gradient must equal the sum of the per-firing gradients (FP64 to
1e-10; the c_ref term is summed over all firings in one reduction) with the
losses adding up.  Firings differ in source position, signal length and
sensor mask size.
"""
import copy

import numpy as np
import pytest

import kwave_amd as K

from cases import rel_l2

pytestmark = pytest.mark.gpu


def _medium(kind, shape):
    rng = np.random.default_rng(3)
    if kind == "het_absorbing":
        c0 = 1500.0 + 40.0 * rng.standard_normal(shape)
        return K.KWaveMedium(sound_speed=c0, density=1000.0 + 20.0 * rng.random(shape), alpha_coeff=0.75,
                             alpha_power=1.5, alpha_mode="stokes")
    if kind == "stokes_nonlinear":
        return K.KWaveMedium(sound_speed=1520.0, density=1000.0, alpha_coeff=0.5, alpha_power=1.2,
                             alpha_mode="stokes", BonA=6.0)
    c0 = np.full(shape, 1500.0)
    c0[shape[0] // 2:, :] = 1580.0
    return K.KWaveMedium(sound_speed=c0, density=1000.0)


def _firings(nx, ny, n, kind):
    kg = K.KWaveGrid(nx, 1e-4, ny, 1e-4)
    med = _medium(kind, (nx, ny))
    K.make_time(kg, 1700.0)
    kg.Nt = 90
    kg.t_array = np.arange(kg.Nt) * kg.dt
    srcs, sens = [], []
    for i in range(n):
        m = np.zeros((nx, ny), dtype=bool)
        m[12 + 3 * i, 10 + 5 * i] = True
        m[12 + 3 * i, 11 + 5 * i] = True
        sig = 1e5 * K.tone_burst(1 / kg.dt, 1.5e6, 2 + (i % 2))
        srcs.append(K.KWaveSource(p_mask=m, p=np.repeat(sig[None, :], 2, axis=0),
                                  p_mode=K.Dirichlet if kind == "dirichlet" else K.Additive))
        sm = np.zeros((nx, ny), dtype=bool)
        sm[nx - 14, 6:ny - 6 - 2 * i] = True
        sens.append(K.KWaveSensor(mask=sm, record=["p", "p_max", "p_rms", "p_final"]))
    return kg, med, srcs, sens


@pytest.mark.parametrize("kind", ["het_absorbing", "stokes_nonlinear", "dirichlet"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_stacked_forward_matches_single_runs(kind, dtype):
    print(f"\n[case] stacked forward, 5 firings 64x48, {kind}, {dtype.__name__}", flush=True)
    kg, med, srcs, sens = _firings(64, 48, 5, kind)
    kw = dict(pml_size=8, pml_alpha=2.0, data_cast=dtype)
    outs = K.kspace_first_order_firings(copy.deepcopy(kg), med, srcs, sens, **kw)
    tol = 1e-13 if dtype == np.float64 else 1e-6
    for i, (s, r) in enumerate(zip(srcs, sens)):
        ref = K.kspace_first_order(copy.deepcopy(kg), med, s, r, smooth_p0=False, **kw)
        for f in ("p", "p_max", "p_rms", "p_final"):
            assert outs[i][f].shape == ref[f].shape, (i, f)
            e = rel_l2(outs[i][f], ref[f])
            assert e <= tol, (i, f, e)
    print("    all firings match", flush=True)


def test_stacked_gradient_matches_sum_of_firings():
    print("\n[case] stacked adjoint gradient, 4 firings 64x48, FP64", flush=True)
    kg, med_true, srcs, sens = _firings(64, 48, 4, "plain")
    for s in sens:
        s.record = ["p"]
    kw = dict(pml_size=8, pml_alpha=2.0, smooth_p0=False)
    obs = [K.kspace_first_order(copy.deepcopy(kg), med_true, s, r, data_cast=np.float64, **kw)["p"]
           for s, r in zip(srcs, sens)]
    c_start = np.full((64, 48), 1500.0) + 3.0 * np.random.default_rng(9).standard_normal((64, 48))
    med0 = K.KWaveMedium(sound_speed=c_start, density=1000.0)
    loss_s, g_s = K.gradient_c0(copy.deepcopy(kg), med0, srcs, sens, obs, checkpoint_every=20,
                                data_cast=np.float64, **kw)
    loss_1, g_1 = K.gradient_c0(copy.deepcopy(kg), med0, srcs, sens, obs, checkpoint_every=20,
                                data_cast=np.float64, stack=1, **kw)
    e = rel_l2(g_s, g_1)
    print(f"    loss {loss_s:.6e} vs {loss_1:.6e}; grad rel-L2 {e:.3e}", flush=True)
    assert loss_s == pytest.approx(loss_1, rel=1e-12)
    assert e <= 1e-10
    # groups of 3 + 1 firings give the same sum
    loss_g, g_g = K.gradient_c0(copy.deepcopy(kg), med0, srcs, sens, obs, checkpoint_every=20,
                                data_cast=np.float64, stack=3, **kw)
    assert loss_g == pytest.approx(loss_1, rel=1e-12)
    assert rel_l2(g_g, g_1) <= 1e-10


def test_stacked_rejects_mismatched_firings():
    kg, med, srcs, sens = _firings(64, 48, 2, "plain")
    p0 = np.zeros((64, 48))
    p0[30, 20] = 1.0
    with pytest.raises(ValueError):
        K.kspace_first_order_firings(copy.deepcopy(kg), med, [K.KWaveSource(p0=p0)], sens[:1])
    from kwave_amd import _lib
    d0, k0, _ = K.build_desc(copy.deepcopy(kg), med, srcs[0], sens[0], pml_size=8)
    med2 = K.KWaveMedium(sound_speed=med.sound_speed * 1.01, density=1000.0)
    d1, k1, _ = K.build_desc(copy.deepcopy(kg), med2, srcs[1], sens[1], pml_size=8)
    with pytest.raises(_lib.KWaveHIPError, match="share the grid"):
        K.Simulation.firings([d0, d1], [k0, k1])
