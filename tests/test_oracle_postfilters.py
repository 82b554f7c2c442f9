"""Pins of the oracle's sensor post-filters (first_order_steps.jl:836-910) by
what a frequency-domain filter of the recorded traces must do:

- frequency response: a DFT-bin tone at the centre frequency passes with
  gain 1, tones at fc +- bw come out scaled by exp(-1/2) (a Gaussian of
  standard deviation bw in |f|), the gain is even in f (real traces stay
  real, a tone keeps its phase), and the response is applied per trace;
- directivity: angle 0 is the identity, a tone is scaled by the
  directivity factor (even in f, 1 at DC), and the factor has its first zero
  where the aperture phase reaches pi; per-sensor angles act per trace.

Planted mistakes (sigma = bw/2, a response in f instead of |f|, the
directivity argument without the 1/2) are rejected.
"""
import math

import numpy as np

from oracle import kwave_oracle as O


class _Sensor:
    def __init__(self, **kw):
        self.directivity_angle = kw.get("directivity_angle")
        self.directivity_size = kw.get("directivity_size")
        self.frequency_response = kw.get("frequency_response")


def _tone(Nt, m, phase=0.0):
    t = np.arange(Nt)
    return np.cos(2 * np.pi * m * t / Nt + phase)


def test_frequency_response_gaussian_in_abs_f():
    Nt, dt = 200, 1e-8
    g = O.Grid((32,), (1e-4,), dt, Nt)
    df = 1.0 / (Nt * dt)                 # 0.5 MHz per bin
    fc, bw = 20 * df, 4 * df
    rows = np.stack([_tone(Nt, 20, 0.3), _tone(Nt, 24), _tone(Nt, 16, -1.0), _tone(Nt, 40), np.ones(Nt)])
    data = {"p": rows.copy()}
    O.apply_frequency_response(data, _Sensor(frequency_response=(fc, bw)), g)
    out = data["p"]
    assert np.allclose(out[0], rows[0], atol=1e-12)                      # centre: gain 1, phase kept
    assert np.allclose(out[1], math.exp(-0.5) * rows[1], atol=1e-12)     # fc + bw
    assert np.allclose(out[2], math.exp(-0.5) * rows[2], atol=1e-12)     # fc - bw
    assert np.allclose(out[3], math.exp(-12.5) * rows[3], atol=1e-12)    # fc + 5 bw
    assert np.allclose(out[4], math.exp(-12.5) * rows[4], atol=1e-12)    # DC: 5 bw below fc
    # planted mistakes: sigma = bw/2, response in signed f
    f = O._fftfreq(Nt, 1 / dt)
    for resp in (np.exp(-((np.abs(f) - fc) ** 2) / (2 * (bw / 2) ** 2)), np.exp(-((f - fc) ** 2) / (2 * bw ** 2))):
        bad = np.fft.ifft(np.fft.fft(rows[1]) * resp).real
        assert np.abs(bad - out[1]).max() > 1e-3


def test_directivity_factor():
    Nt, dt = 128, 2e-8
    g = O.Grid((40, 30), (1e-4, 2e-4), dt, Nt)
    dmin = 1e-4
    df = 1.0 / (Nt * dt)
    m = 8
    k = 2 * math.pi * m * df / dmin
    size = 3e-4
    a_zero = math.asin(2 * math.pi / (k * size))         # aperture phase k size sin(a) / 2 = pi
    a_one = math.asin(2 * 1.2 / (k * size))              # aperture phase 1.2 rad
    angles = np.array([0.0, a_zero, a_one])
    rows = np.stack([_tone(Nt, m), _tone(Nt, m, 0.7), _tone(Nt, m)])
    data = {"p": rows.copy()}
    O.apply_sensor_directivity(data, _Sensor(directivity_angle=angles, directivity_size=size), g)
    out = data["p"]
    assert np.array_equal(out[0], rows[0])                                # angle 0: untouched
    assert np.abs(out[1]).max() < 1e-12                                  # first zero of the factor
    x = k * size * math.sin(a_one) / 2
    assert np.allclose(out[2], math.sin(x) / x * rows[2], atol=1e-12)     # even in f: a scaled tone
    dc = {"p": np.ones((1, Nt))}
    O.apply_sensor_directivity(dc, _Sensor(directivity_angle=a_one, directivity_size=size), g)
    assert np.allclose(dc["p"], 1.0, atol=1e-12)                          # DC passes
    # planted mistake: no 1/2 in the aperture phase
    assert np.abs(math.sin(2 * x) / (2 * x) * rows[2] - out[2]).max() > 1e-3
