"""GPU runs driven by an off-grid transducer array (array.jl): one firing per
element through the batch form, element signals recovered with
combine_sensor_data, against the oracle on the same distributed inputs; and
a focused, steered, apodized linear-transducer firing (transducer.jl)."""
import copy

import numpy as np
import pytest

import kwave_amd as K
from oracle import kwave_oracle as O

from cases import rel_l2

pytestmark = pytest.mark.gpu


def test_array_firings_batch_matches_oracle():
    print("\n[case] transducer-array firings (2 arcs), batch f64", flush=True)
    kg = K.KWaveGrid(96, 1e-4, 96, 1e-4)
    c0 = 1500.0 + 60.0 * np.random.default_rng(5).random((96, 96))
    K.make_time(kg, c0)
    kg.Nt = 80
    kg.t_array = np.arange(80) * kg.dt
    med = K.KWaveMedium(sound_speed=c0, density=1000.0, alpha_coeff=0.5, alpha_power=1.1, alpha_mode="stokes")
    arr = K.KWaveArray()
    K.add_arc_element(arr, (0.0, 0.0), 25e-4, 18e-4, (25e-4, 0.0))
    K.add_arc_element(arr, (0.0, 0.0), 25e-4, 18e-4, (-25e-4, 0.0))
    burst = K.tone_burst(1 / kg.dt, 2e6, 2)
    sims = []
    for e in range(2):
        sig = np.zeros((2, kg.Nt))
        sig[e, : burst.size] = burst
        mask, dist = K.get_distributed_source_signal(arr, kg, sig)
        src = K.KWaveSource(p_mask=mask, p=dist)
        sen = K.KWaveSensor(mask=K.get_array_binary_mask(arr, kg), record=["p"])
        sims.append((kg, med, src, sen))
    kw = dict(pml_size=12, pml_alpha=2.0, smooth_p0=False)
    outs = K.kspace_first_order_batch([tuple(copy.deepcopy(s)) for s in sims], data_cast=np.float64, **kw)
    for e, (kg_, med_, src, sen) in enumerate(sims):
        ref = O.kspace_first_order(copy.deepcopy(kg_), med_, src, sen, **kw)
        assert rel_l2(outs[e]["p"], ref["p"]) <= 1e-10
        per_elem = K.combine_sensor_data(arr, kg, outs[e]["p"])
        ref_elem = K.combine_sensor_data(arr, kg, ref["p"])
        assert per_elem.shape == (2, kg.Nt)
        assert rel_l2(per_elem, ref_elem) <= 1e-10
        # the firing element hears its own pulse first and loudest
        assert np.abs(per_elem[e]).max() > np.abs(per_elem[1 - e]).max()


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_transducer_firing_matches_oracle(dtype):
    print(f"\n[case] linear transducer firing 3-D {dtype.__name__}", flush=True)
    kg = K.KWaveGrid(32, 1e-4, 64, 1e-4, 32, 1e-4)
    K.make_time(kg, 1540.0)
    kg.Nt = 60
    kg.t_array = np.arange(60) * kg.dt
    med = K.KWaveMedium(sound_speed=1540.0, density=1000.0)
    tr = K.KWaveTransducer(number_elements=8, element_width=3, element_spacing=1, element_length=6,
                           position=(6, 8, 12), focus_distance=2e-3, steering_angle=0.1,
                           transmit_apodization="hann", receive_apodization="hamming",
                           input_signal=1e5 * K.tone_burst(1 / kg.dt, 2e6, 2))
    mask, sig = K.get_transducer_source(tr, kg, med)
    src = K.KWaveSource(p_mask=mask, p=sig)
    sen = K.KWaveSensor(mask=mask, record=["p"])
    kw = dict(pml_size=6, pml_alpha=2.0, smooth_p0=False)
    out = K.kspace_first_order(copy.deepcopy(kg), med, src, sen, data_cast=dtype, **kw)
    ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
    tol = 1e-10 if dtype == np.float64 else 1e-4
    assert rel_l2(out["p"], ref["p"]) <= tol
    got = K.combine_transducer_sensor_data(tr, kg, out["p"])
    want = K.combine_transducer_sensor_data(tr, kg, ref["p"])
    assert got.shape == (8, 60) and rel_l2(got, want) <= tol
