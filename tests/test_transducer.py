"""Linear-array transducer (transducer.jl): the reference's assertions
(test_transducer.jl) plus what the definitions fix -- apodization equals the
textbook symmetric windows (numpy), focusing delays follow the geometric path
difference and steering delays are linear in the element position, each
element's rows carry the input signal shifted by round(delay/dt), and
combining the transmit rows gives the per-element signal times both
apodizations."""
import numpy as np
import pytest

import kwave_amd as K
from kwave_amd.transducer import compute_apodization, transmit_delays


def _kg():
    return K.KWaveGrid(32, 1e-4, 64, 1e-4, 32, 1e-4)


def test_reference_assertions():
    t = K.KWaveTransducer(number_elements=32, element_width=2, element_length=10)
    assert t.number_elements == 32 and t.element_width == 2 and t.element_spacing == 0
    assert t.radius == float("inf") and t.transmit_apodization == "rectangular"
    kg = _kg()
    t = K.KWaveTransducer(number_elements=8, element_width=4, element_length=16, position=(16, 5, 8))
    m = K.get_transducer_binary_mask(t, kg)
    assert m.shape == (32, 64, 32) and m.sum() == 8 * 4 * 16
    t = K.KWaveTransducer(number_elements=8, element_width=4, element_length=16, position=(16, 5, 8),
                          active_elements=[1, 3, 5])
    assert K.get_transducer_binary_mask(t, kg).sum() == 3 * 4 * 16
    K.make_time(kg, 1500.0, t_end=2e-6)
    sig = np.sin(2 * np.pi * 1e6 * kg.t_array)
    t = K.KWaveTransducer(number_elements=4, element_width=2, element_length=8, position=(16, 10, 12), input_signal=sig)
    mask, src = K.get_transducer_source(t, kg, K.KWaveMedium(sound_speed=1500.0))
    assert mask.sum() > 0 and src.shape == (mask.sum(), kg.Nt)
    with pytest.raises(TypeError):
        K.get_transducer_binary_mask(t, K.KWaveGrid(32, 1e-4, 32, 1e-4))


@pytest.mark.parametrize("kind,ref", [("hann", np.hanning), ("hanning", np.hanning), ("hamming", np.hamming),
                                      ("blackman", np.blackman), ("bartlett", np.bartlett),
                                      ("triangular", np.bartlett)])
def test_apodization_windows(kind, ref):
    for n in (2, 5, 16, 33):
        assert np.allclose(compute_apodization(n, kind), ref(n), rtol=0, atol=1e-14)
    assert np.array_equal(compute_apodization(7, "rectangular"), np.ones(7))
    with pytest.raises(ValueError):
        compute_apodization(4, "kaiserish")


def test_delays_focus_and_steering():
    kg = _kg()
    c0 = 1540.0
    t = K.KWaveTransducer(number_elements=9, element_width=3, element_spacing=1, element_length=4,
                          position=(5, 2, 3), focus_distance=2e-3)
    d = transmit_delays(t, kg, c0)
    pitch = 4 * kg.dy
    y = (np.arange(9) - 4) * pitch
    want = (np.hypot(2e-3, y) - 2e-3) / c0
    assert np.allclose(d, want - want.min(), rtol=0, atol=1e-18)
    # the reference delays each element by its extra path: the centre fires first
    assert d[4] == 0.0 and d[0] == d.max() and np.isclose(d[0], d[8], rtol=0, atol=1e-20)
    s = K.KWaveTransducer(number_elements=6, element_width=2, element_length=2, position=(5, 2, 3),
                          steering_angle=0.3)
    d = transmit_delays(s, kg, c0)
    assert np.allclose(np.diff(d), 2 * kg.dy * np.sin(0.3) / c0, rtol=1e-12)
    assert d.min() == 0.0


def test_signal_shift_rows_and_combination():
    kg = _kg()
    K.make_time(kg, 1500.0, t_end=4e-6)
    sig = np.sin(2 * np.pi * 2e6 * kg.t_array[:30])
    t = K.KWaveTransducer(number_elements=5, element_width=2, element_spacing=1, element_length=3,
                          position=(8, 4, 6), focus_distance=1.5e-3, transmit_apodization="hann",
                          receive_apodization="hamming", input_signal=sig)
    mask, src = K.get_transducer_source(t, kg, K.KWaveMedium(sound_speed=1500.0))
    assert src.shape == (5 * 2 * 3, kg.Nt) and mask.sum() == 30
    d = transmit_delays(t, kg, 1500.0)
    apod = np.hanning(5)
    for e in range(5):
        rows = src[6 * e:6 * e + 6]
        shift = int(round(d[e] / kg.dt))
        want = np.zeros(kg.Nt)
        want[shift:shift + 30] = sig * apod[e]
        assert np.allclose(rows, np.repeat(want[None, :], 6, axis=0), rtol=0, atol=1e-15), e
        assert np.array_equal(rows != 0, np.repeat((want != 0)[None, :], 6, axis=0)), e
    comb = K.combine_transducer_sensor_data(t, kg, src)
    for e in range(5):
        shift = int(round(d[e] / kg.dt))
        want = np.zeros(kg.Nt)
        want[shift:shift + 30] = sig * apod[e] * np.hamming(5)[e]
        assert np.allclose(comb[e], want, rtol=0, atol=1e-15)
    # rows follow the element loop order (z outer, y inner), not findall order
    t2 = K.KWaveTransducer(number_elements=2, element_width=2, element_length=2, position=(3, 1, 1))
    pts = []
    for e in (1, 2):
        pts += [(3, j, k) for k in (1, 2) for j in ((1, 2) if e == 1 else (3, 4))]
    lin = [(i - 1) + kg.Nx * ((j - 1) + kg.Ny * (k - 1)) for i, j, k in pts]
    assert lin != sorted(lin)                       # the two orders differ for this layout
    m2 = K.get_transducer_binary_mask(t2, kg)
    assert sorted(lin) == list(np.flatnonzero(m2.ravel(order="F")))
