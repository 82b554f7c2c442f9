"""Operator-level GPU parity: one spectral derivative along one axis
(spectral_gradient!, fft_utils.jl:126-182) through kw_test_spectral_gradient,
against the analytic sinusoid of test_fft.jl:24-56 and the oracle's N-D
rfft restatement on random fields (odd and even lengths, every axis, both
staggered directions)."""
import numpy as np
import pytest

import kwave_amd as K
from kwave_amd import operators as OP
from oracle import kwave_oracle as O

pytestmark = pytest.mark.gpu


def test_reference_sinusoid_64x64():
    print("\n[case] spectral gradient sin(kx) 64x64 f64", flush=True)
    Nx = Ny = 64
    x = O.make_spatial_vec(Nx, 1.0)
    k = 2 * np.pi / Nx
    f = np.repeat(np.sin(k * x)[:, None], Ny, axis=1)
    df = OP.spectral_gradient(f, (1.0, 1.0), axis=0, sign=0)
    want = k * np.cos(k * x)
    assert np.allclose(df[4:Nx - 4, Ny // 2], want[4:Nx - 4], rtol=0, atol=1e-10)


SHAPES = [(64,), (45,), (32, 24), (27, 40), (16, 25, 12), (20, 9, 15)]


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_spectral_gradient_matches_oracle(shape, dtype):
    print(f"\n[case] spectral gradient {shape} {dtype.__name__}", flush=True)
    rng = np.random.default_rng(len(shape) * 100 + shape[0])
    f = rng.standard_normal(shape)
    d = [1e-4 * (1 + 0.2 * a) for a in range(len(shape))]
    tol = 1e-12 if dtype == np.float64 else 2e-6
    for axis in range(len(shape)):
        kv = O.make_wavenumber_vec(shape[axis], d[axis])
        sp, sm = O.make_shift_operators(kv, d[axis])
        signs = [(1, sp), (-1, sm)]
        if axis == 0 or shape[axis] % 2 == 1:
            signs.append((0, np.ones_like(sp)))
        for sign, shift in signs:
            want = O.spectral_gradient(f, kv, shift, axis)
            got = OP.spectral_gradient(f, d, axis, sign, data_cast=dtype)
            e = np.abs(got - want).max() / np.abs(want).max()
            print(f"    axis {axis} sign {sign:+d} max rel err {e:.2e}", flush=True)
            assert e <= tol, (axis, sign, e)


def test_unseparable_operator_is_refused():
    print("\n[case] unstaggered y derivative of an even axis refused", flush=True)
    with pytest.raises(K.KWaveHIPError):
        OP.spectral_gradient(np.zeros((8, 8)), (1e-4, 1e-4), axis=1, sign=0)
