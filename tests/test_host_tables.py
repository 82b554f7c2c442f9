"""Operator-level parity of the host tables the C++ side builds (through the
C ABI, no GPU): wavenumbers, PML profiles, staggered derivative multipliers
and EOS multipliers against the oracle's (pinned) restatements."""
import copy

import numpy as np
import pytest

import kwave_amd as K
from kwave_amd import operators as OP
from oracle import kwave_oracle as O


@pytest.mark.parametrize("n", [7, 8, 9, 15, 64, 128, 135])
def test_axis_tables(n):
    d, dt = 1.3e-4, 2e-8
    k = OP.axis_table(n, d, "k")
    assert np.allclose(k, O.make_wavenumber_vec(n, d), rtol=1e-15, atol=0)
    for size, alpha in ((0, 2.0), (3, 2.0), (min(10, n), 1e7)):
        p = OP.axis_table(n, d, "pml", pml_size=size, pml_alpha=alpha, dt=dt)
        ps = OP.axis_table(n, d, "pml_sg", pml_size=size, pml_alpha=alpha, dt=dt)
        assert np.allclose(p, O.get_pml(n, d, size, alpha, dt), rtol=1e-15, atol=0)
        assert np.allclose(ps, O.get_pml(n, d, size, alpha, dt, staggered=True), rtol=1e-15, atol=0)
    kv = O.make_wavenumber_vec(n, d)
    sp, sm = O.make_shift_operators(kv, d)
    for which, s in (("dplus", sp), ("dminus", sm)):
        want = 1j * kv * s
        if n % 2 == 0:
            want[n // 2] = want[n // 2].real          # what the c2r inverse keeps at Nyquist
        got = OP.axis_table(n, d, which)
        assert np.allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


def _desc(shape, medium, c_ref=None):
    kg = K.KWaveGrid(*sum(([n, 1e-4 * (1 + 0.1 * i)] for i, n in enumerate(shape)), []))
    K.make_time(kg, 1500.0)
    sen = K.KWaveSensor(mask=np.ones(shape, dtype=bool), record=["p"])
    src = K.KWaveSource(p0=np.ones(shape))
    d, keep, info = K.build_desc(kg, medium, src, sen, smooth_p0=False, pml_size=2)
    return kg, d, keep


CASES = [
    ("lossless", dict(sound_speed=1500.0, density=1000.0)),
    ("no_dispersion", dict(sound_speed=1480.0, density=1000.0, alpha_coeff=0.75, alpha_power=1.5,
                           alpha_mode="no_dispersion")),
    ("stokes", dict(sound_speed=1520.0, density=1000.0, alpha_coeff=0.5, alpha_power=1.1, alpha_mode="stokes")),
]


@pytest.mark.parametrize("shape", [(32,), (9, 16), (8, 15, 6)])
@pytest.mark.parametrize("name,mkw", CASES, ids=[c[0] for c in CASES])
def test_eos_tables(shape, name, mkw):
    med = K.KWaveMedium(**mkw)
    kg, d, keep = _desc(shape, med)
    g = O.grid_from(kg)
    c_ref = float(mkw["sound_speed"])
    kap = O.compute_kappa(g.k, c_ref, g.dt)
    ab = O.precompute_absorption(med, g.k, c_ref)
    m1 = c_ref ** 2 * kap * (1 + (ab["nabla1"] if ab is not None else 0.0))
    nh = shape[0] // 2 + 1
    got = OP.eos_table(d, 0)
    assert got.shape == (nh,) + tuple(shape[1:])
    assert np.allclose(got, m1[:nh], rtol=1e-13, atol=0)
    if name == "stokes":
        with np.errstate(divide="ignore", invalid="ignore"):
            m2 = -c_ref ** 2 * kap * ab["nabla2"]
        m2 = np.where(g.k == 0, 0.0, m2)[:nh]       # |k|^(y-1) at k = 0: the kernel multiplies a zero DC bin
        got2 = OP.eos_table(d, 1)
        fin = g.k[:nh] != 0
        assert np.allclose(got2[fin], m2[fin], rtol=1e-13, atol=0)
    else:
        with pytest.raises(K.KWaveHIPError):
            OP.eos_table(d, 1)


def test_eos_table_heterogeneous_c0():
    shape = (12, 10)
    c0 = 1500.0 + 40.0 * np.random.default_rng(1).random(shape)
    med = K.KWaveMedium(sound_speed=c0, density=1000.0, alpha_coeff=0.75, alpha_power=1.5, alpha_mode="stokes")
    kg, d, keep = _desc(shape, med)
    g = O.grid_from(kg)
    c_ref = float(np.mean(c0))
    assert abs(d.c_ref - c_ref) <= 1e-12 * c_ref
    kap = O.compute_kappa(g.k, d.c_ref, g.dt)
    ab = O.precompute_absorption(med, g.k, d.c_ref)
    want = (kap * (1 + ab["nabla1"]))[: shape[0] // 2 + 1]       # c0^2 applied in real space, no dispersion
    assert np.allclose(OP.eos_table(d, 0), want, rtol=1e-13, atol=0)
