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

import pytest

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


def test_frequency_response_law_at_every_bin():
    """Every DFT bin is an eigenvector (a tone keeps its phase) and log(gain) is one
    quadratic in |f| with its vertex at fc and curvature -1/bw^2: the Gaussian law
    holds across the whole band, not only at fc and fc +- bw."""
    Nt, dt = 96, 1e-8
    g = O.Grid((16,), (1e-4,), dt, Nt)
    df = 1.0 / (Nt * dt)
    fc, bw = 13.5 * df, 6.25 * df                # centre between bins
    ms = np.arange(0, Nt // 2)
    rows = np.stack([_tone(Nt, m, 0.1 * m) for m in ms])
    data = {"p": rows.copy()}
    O.apply_frequency_response(data, _Sensor(frequency_response=(fc, bw)), g)
    gains = []
    for m, r, o in zip(ms, rows, data["p"]):
        gm = float(np.dot(o, r) / np.dot(r, r))
        assert np.allclose(o, gm * r, atol=1e-13), m             # eigenvector, phase kept
        gains.append(gm)
    lg = np.log(np.asarray(gains))
    f = ms * df
    # log G = -(f - fc)^2 / (2 bw^2): second difference constant, first difference zero at fc
    second = lg[2:] - 2 * lg[1:-1] + lg[:-2]
    assert np.allclose(second, -(df / bw) ** 2, rtol=0, atol=1e-9)
    first = np.diff(lg) / df                                     # slope at the midpoints
    mid = (f[1:] + f[:-1]) / 2
    assert np.allclose(first, -(mid - fc) / bw ** 2, rtol=1e-9, atol=1e-18)
    assert lg.max() <= 0 and np.argmax(lg) in (13, 14)


def test_directivity_law_at_every_bin():
    """Every bin is scaled by sin(c m)/(c m) with one constant c per sensor (the
    aperture phase grows linearly with |f|), and the sensor with angle 0 is untouched."""
    Nt, dt = 80, 2e-8
    g = O.Grid((30, 30), (1e-4, 1e-4), dt, Nt)
    df = 1.0 / (Nt * dt)
    size, ang = 4e-4, 0.3
    c = (2 * math.pi * df / 1e-4) * size * math.sin(ang) / 2      # aperture phase per bin
    for m in range(1, Nt // 2):
        data = {"p": np.stack([_tone(Nt, m, 0.2), _tone(Nt, m, 0.2)])}
        ref = data["p"].copy()
        O.apply_sensor_directivity(data, _Sensor(directivity_angle=np.array([ang, 0.0]), directivity_size=size), g)
        gm = float(np.dot(data["p"][0], ref[0]) / np.dot(ref[0], ref[0]))
        assert np.allclose(data["p"][0], gm * ref[0], atol=1e-13)
        assert gm == pytest.approx(math.sin(c * m) / (c * m), rel=1e-10, abs=1e-13)
        assert np.array_equal(data["p"][1], ref[1])
