"""Parity at the benchmark workloads' full grid sizes, through properties that
do not need the oracle to run at that size (it finishes 256^3 steps in
seconds each, 512^3 not at all):

- the FP32 kernels against the FP64 build of the same kernels (itself
  parity-tested against the oracle at small sizes to 1e-10) on the 256^3
  heterogeneous absorbing nonlinear workload, within the 1e-4 rel-L2 bar;
- superposition on the 256^3 grid with the medium made linear and lossless:
  the response to two sources driven together equals the sum of the
  responses to each alone (FP64, to rounding);
- the 512^3 FP16-map workload: FP32 arithmetic on FP16 maps against FP64
  arithmetic on the same dequantised maps, within 1e-4.
Each run is 150-200 steps (30 at 512^3), enough for the waves to cross the
recorded points.
"""
import copy

import numpy as np
import pytest

import kwave_amd as K
from kwave_amd.configs import ellipsoids_3d, layered_3d_fp16

from cases import rel_l2

pytestmark = pytest.mark.gpu


def _near_plane_sensor(N, z, record):
    m = np.zeros((N, N, N), dtype=bool)
    m[:, :, z] = True
    return K.KWaveSensor(mask=m, record=record)


def test_c2_fp32_vs_fp64_256cube():
    print("\n[case] 256^3 heterogeneous absorbing nonlinear, FP32 vs FP64, 200 steps", flush=True)
    kg, med, src, _, kw = ellipsoids_3d(N=256, Nt=200)
    kw.pop("data_cast")
    sen = _near_plane_sensor(256, 60, ["p", "p_max", "p_min", "p_rms", "p_final"])
    o64 = K.kspace_first_order(copy.deepcopy(kg), med, src, sen, data_cast=np.float64, **kw)
    o32 = K.kspace_first_order(copy.deepcopy(kg), med, src, sen, data_cast=np.float32, **kw)
    assert np.abs(o64["p_max"]).max() > 1e5          # the pulse crossed the plane
    for k in ("p", "p_max", "p_min", "p_rms", "p_final"):
        e = rel_l2(o32[k], o64[k])
        print(f"    {k:8s} rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-4, (k, e)


def test_superposition_256cube():
    print("\n[case] 256^3 linear lossless superposition, FP64, 150 steps", flush=True)
    kg, med, src, _, kw = ellipsoids_3d(N=256, Nt=150)
    kw.pop("data_cast")
    lin = K.KWaveMedium(sound_speed=med.sound_speed, density=med.density)
    pm = src.p_mask
    n = int(pm.sum())
    rows = np.arange(n)
    half = rows < n // 2
    sig = src.p
    shifted = np.zeros_like(sig)
    shifted[:, 7:] = 0.6 * sig[:, :-7]
    sA = np.where(half[:, None], sig, 0.0)
    sB = np.where(half[:, None], 0.0, shifted)
    sen = _near_plane_sensor(256, 50, ["p", "p_final"])
    run = lambda s: K.kspace_first_order(copy.deepcopy(kg), lin, K.KWaveSource(p_mask=pm, p=s), sen,
                                         data_cast=np.float64, **kw)
    a, b, ab = run(sA), run(sB), run(sA + sB)
    for k in ("p", "p_final"):
        e = rel_l2(a[k] + b[k], ab[k])
        print(f"    {k:8s} superposition rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-11, (k, e)
    assert rel_l2(a["p"], ab["p"]) > 0.1              # both sources matter


def test_c4_fp16_maps_512cube():
    print("\n[case] 512^3 FP16 maps: FP32 arithmetic vs FP64 on the dequantised maps, 30 steps", flush=True)
    kg, med, src, sensors, kw = layered_3d_fp16(Nt=30)
    kw = dict(kw)
    kw.pop("medium_storage")
    sen = K.KWaveSensor(mask=src.p_mask, record=["p", "p_final"])
    o16 = K.kspace_first_order(copy.deepcopy(kg), med, src, sen, data_cast=np.float32,
                               medium_storage=np.float16, **kw)
    cq, _ = K.quantize_map(med.sound_speed)
    rq, _ = K.quantize_map(med.density)
    medq = K.KWaveMedium(sound_speed=cq, density=rq, alpha_coeff=med.alpha_coeff, alpha_power=med.alpha_power,
                         alpha_mode=med.alpha_mode, BonA=med.BonA)
    del cq, rq
    o64 = K.kspace_first_order(copy.deepcopy(kg), medq, src, sen, data_cast=np.float64, **kw)
    for k in ("p", "p_final"):
        e = rel_l2(o16[k], o64[k])
        print(f"    {k:8s} rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-4, (k, e)
