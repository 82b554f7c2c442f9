"""Axis lengths without a compiled line-FFT plan: chirp-z (Bluestein) lines, against the oracle.

The reference takes any grid size (FFTW plans any length, fft_utils.jl:44-62;
odd N in grid.jl:126-128).  Lengths without a compiled plan run their line
transforms as chirp-z convolutions on a power-of-two plan of at least 2n - 1
points.  Every case runs the whole simulation on the GPU through the C ABI and
compares every recorded array with the oracle (scipy.fft, any length): FP64
kernels to 1e-10 and FP32 kernels to 1e-4 relative L2, as for compiled
lengths.  Primes, odd and even composites, mixed compiled/chirp-z axes, both
EOS spectra (homogeneous :stokes), p0 with smoothing, sources and sensors.
"""
import copy

import numpy as np
import pytest

import kwave_amd as K
from kwave_amd import _lib
from oracle import kwave_oracle as O

from cases import rel_l2

pytestmark = pytest.mark.gpu


def _grid(shape, d=1e-4):
    args = []
    for n in shape:
        args += [n, d]
    return K.KWaveGrid(*args)


def _problem(shape, kind, seed):
    rng = np.random.default_rng(seed)
    nd = len(shape)
    kg = _grid(shape)
    if kind == "het_absorb_nonlinear":
        c0 = 1500.0 + 80.0 * rng.random(shape)
        med = K.KWaveMedium(sound_speed=c0, density=1000.0 + 50.0 * rng.random(shape), alpha_coeff=0.75,
                            alpha_power=1.5, alpha_mode="stokes", BonA=5.0)
    elif kind == "hom_stokes":
        med = K.KWaveMedium(sound_speed=1540.0, density=1000.0, alpha_coeff=1.0, alpha_power=1.3,
                            alpha_mode="stokes", BonA=4.0)
    else:
        med = K.KWaveMedium(sound_speed=1500.0, density=1000.0)
    K.make_time(kg, 1700.0)
    kg.Nt = 70
    kg.t_array = np.arange(kg.Nt) * kg.dt
    n = int(np.prod(shape))
    m = np.zeros(n, dtype=bool)
    m[rng.choice(n, size=min(5, n // 4), replace=False)] = True
    m = m.reshape(shape, order="F")
    sig = 1e5 * K.tone_burst(1 / kg.dt, 1.5e6, 2)
    src = K.KWaveSource(p_mask=m, p=np.repeat(sig[None, :], int(m.sum()), axis=0))
    sm = np.zeros(n, dtype=bool)
    sm[rng.choice(n, size=min(40, n // 3), replace=False)] = True
    sen = K.KWaveSensor(mask=sm.reshape(shape, order="F"), record=["p", "p_max", "p_rms", "p_final", "ux"])
    return kg, med, src, sen


CASES = [
    ((7,), "plain"), ((101,), "het_absorb_nonlinear"), ((257,), "hom_stokes"), ((1000,), "plain"),
    ((37, 50), "het_absorb_nonlinear"), ((64, 37), "hom_stokes"), ((130, 11), "plain"),
    ((21, 17, 26), "het_absorb_nonlinear"), ((16, 13, 24), "hom_stokes"),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("shape,kind", CASES, ids=[f"{'x'.join(map(str, c[0]))}-{c[1]}" for c in CASES])
def test_chirp_z_axes_match_oracle(shape, kind, dtype):
    print(f"\n[case] chirp-z axes {shape} {kind} {dtype.__name__}", flush=True)
    lib = _lib.load()
    assert any(lib.kw_fft_length_supported(n) == 2 for n in shape)
    kg, med, src, sen = _problem(shape, kind, sum(shape))
    kw = dict(pml_size=4, pml_alpha=2.0, smooth_p0=False)
    ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
    out = K.kspace_first_order(copy.deepcopy(kg), med, src, sen, data_cast=dtype, **kw)
    tol = 1e-10 if dtype == np.float64 else 1e-4
    for k in ("p", "p_max", "p_rms", "p_final", "ux"):
        e = rel_l2(out[k], ref[k])
        print(f"    {k:8s} rel-L2 {e:.3e}", flush=True)
        assert np.abs(ref[k]).max() > 0, k
        assert e <= tol, (k, e)


@pytest.mark.parametrize("shape", [(45, 29), (31, 24, 20)], ids=["45x29", "31x24x20"])
def test_chirp_z_smoothed_p0(shape):
    print(f"\n[case] chirp-z axes {shape}, p0 with smoothing, FP64", flush=True)
    kg = _grid(shape)
    K.make_time(kg, 1500.0)
    kg.Nt = 40
    kg.t_array = np.arange(kg.Nt) * kg.dt
    rng = np.random.default_rng(11)
    p0 = np.zeros(shape)
    sl = tuple(slice(n // 2 - 2, n // 2 + 2) for n in shape)
    p0[sl] = 1.0 + rng.random(p0[sl].shape)
    med = K.KWaveMedium(sound_speed=1500.0 + 30.0 * rng.random(shape), density=1000.0)
    src = K.KWaveSource(p0=p0)
    sen = K.KWaveSensor(mask=np.ones(shape, dtype=bool), record=["p_final", "p_max"])
    kw = dict(pml_size=3, pml_alpha=2.0, smooth_p0=True)
    ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
    out = K.kspace_first_order(copy.deepcopy(kg), med, src, sen, data_cast=np.float64, **kw)
    for k in ("p_final", "p_max"):
        e = rel_l2(out[k], ref[k])
        print(f"    {k:8s} rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-10, (k, e)


def test_chirp_z_lengths_outside_the_limit_are_rejected():
    kg = _grid((3001, 16))
    K.make_time(kg, 1500.0)
    kg.Nt = 2
    med = K.KWaveMedium(sound_speed=1500.0, density=1000.0)
    src = K.KWaveSource(p0=np.ones((3001, 16)))
    sen = K.KWaveSensor(mask=np.ones((3001, 16), dtype=bool), record=["p_final"])
    with pytest.raises(K.KWaveHIPError, match="chirp-z"):
        K.kspace_first_order(kg, med, src, sen, smooth_p0=False)
