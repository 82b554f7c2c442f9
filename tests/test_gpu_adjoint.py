"""GPU parity of the checkpointed adjoint (kw_sim_grad_c0) against the oracle.

FP64 kernels must reproduce the oracle's gradient to round-off; FP32 kernels
within 1e-3 relative L2; the FP64 GPU gradient must also match central
finite differences of the GPU forward run on a few pixels to 1e-3.
"""
import copy

import numpy as np
import pytest

import kwave_amd as K
from oracle import kwave_oracle as O

from cases import rel_l2, scenario

pytestmark = pytest.mark.gpu


def _setup(name, shape, nt, seed=3):
    kg, med, src, sen, kw = scenario(name, shape, seed=seed, nt=nt)
    sen.record = ["p"]
    rng = np.random.default_rng(seed + 100)
    c_true = np.asarray(med.sound_speed, dtype=np.float64) + 30.0 * rng.standard_normal(shape)
    med_true = copy.deepcopy(med)
    med_true.sound_speed = c_true
    obs = O.kspace_first_order(copy.deepcopy(kg), med_true, src, sen, **kw)["p"]
    return kg, med, src, sen, kw, obs


CASES = [
    ("hetc_absorb_psrc", (64, 48), 60, 16),
    ("hetc_hetr_psrc_p0", (48, 40), 45, 10),
    ("hetc_absorb_psrc", (128,), 80, 25),
    ("hetc_hetr_psrc", (16, 20, 12), 30, 7),
    ("hetc_absorb_p0", (24, 16, 20), 24, 24),
]


@pytest.mark.parametrize("name,shape,nt,ck", CASES, ids=[f"{c[0]}-{len(c[1])}d" for c in CASES])
def test_grad_c0_fp64_matches_oracle(name, shape, nt, ck):
    print(f"\n[case] adjoint fp64 {name} {shape} nt={nt} ckpt={ck}", flush=True)
    kg, med, src, sen, kw, obs = _setup(name, shape, nt)
    loss_ref, g_ref = O.grad_c0(copy.deepcopy(kg), med, src, sen, obs, pml_size=kw["pml_size"],
                                pml_alpha=kw["pml_alpha"], smooth_p0=kw["smooth_p0"])
    loss, g = K.gradient_c0(copy.deepcopy(kg), med, src, sen, obs, checkpoint_every=ck, data_cast=np.float64, **kw)
    print(f"    loss {loss:.6e} ref {loss_ref:.6e}; grad rel-L2 {rel_l2(g, g_ref):.3e}", flush=True)
    assert loss == pytest.approx(loss_ref, rel=1e-10)
    assert rel_l2(g, g_ref) <= 1e-9


def test_grad_c0_fp32_matches_oracle():
    print("\n[case] adjoint fp32 2d", flush=True)
    kg, med, src, sen, kw, obs = _setup("hetc_absorb_psrc", (64, 48), 60)
    loss_ref, g_ref = O.grad_c0(copy.deepcopy(kg), med, src, sen, obs, pml_size=kw["pml_size"],
                                pml_alpha=kw["pml_alpha"], smooth_p0=False)
    loss, g = K.gradient_c0(copy.deepcopy(kg), med, src, sen, obs, checkpoint_every=16, data_cast=np.float32, **kw)
    e = rel_l2(g, g_ref)
    print(f"    loss {loss:.6e} ref {loss_ref:.6e}; grad rel-L2 {e:.3e}", flush=True)
    assert loss == pytest.approx(loss_ref, rel=1e-3)
    assert e <= 1e-3


def test_grad_c0_checkpoint_spacing_is_exact():
    """Every checkpoint spacing gives the same gradient (recompute is deterministic)."""
    kg, med, src, sen, kw, obs = _setup("hetc_absorb_psrc", (48, 48), 40)
    gs = [K.gradient_c0(copy.deepcopy(kg), med, src, sen, obs, checkpoint_every=ck, data_cast=np.float32, **kw)[1]
          for ck in (1, 7, 40)]
    assert np.array_equal(gs[0], gs[1]) and np.array_equal(gs[0], gs[2])


def test_grad_c0_multi_firing_sum():
    kg, med, src, sen, kw, obs = _setup("hetc_absorb_psrc", (48, 40), 40, seed=3)
    _, _, src2, _, _, obs2 = _setup("hetc_absorb_psrc", (48, 40), 40, seed=4)
    # the second firing keeps the first one's medium and sensor
    obs2 = O.kspace_first_order(copy.deepcopy(kg), med, src2, sen, **kw)["p"] * 0.5
    l1, g1 = K.gradient_c0(copy.deepcopy(kg), med, src, sen, obs, data_cast=np.float64, **kw)
    l2, g2 = K.gradient_c0(copy.deepcopy(kg), med, src2, sen, obs2, data_cast=np.float64, **kw)
    l, g = K.gradient_c0(copy.deepcopy(kg), med, [src, src2], sen, [obs, obs2], data_cast=np.float64, **kw)
    assert l == pytest.approx(l1 + l2, rel=1e-12)
    assert np.allclose(g, g1 + g2, rtol=1e-12, atol=0)


def test_grad_c0_finite_differences_on_gpu():
    kg, med, src, sen, kw, obs = _setup("hetc_absorb_psrc", (64, 48), 50)
    _, g = K.gradient_c0(copy.deepcopy(kg), med, src, sen, obs, checkpoint_every=10, data_cast=np.float64, **kw)
    c0 = np.asarray(med.sound_speed, dtype=np.float64)

    def loss(c):
        m = copy.deepcopy(med)
        m.sound_speed = c
        p = K.kspace_first_order(copy.deepcopy(kg), m, src, sen, data_cast=np.float64, **kw)["p"]
        return 0.5 * float(np.sum((p - obs) ** 2))

    gf = g.ravel(order="F")
    idx = np.argsort(-np.abs(gf))[:3].tolist() + np.random.default_rng(2).choice(c0.size, 3, replace=False).tolist()
    for i in idx:
        h = 1e-5 * c0.ravel(order="F")[i]
        e = np.zeros(c0.size)
        e[i] = h
        e = e.reshape(c0.shape, order="F")
        fd = (loss(c0 + e) - loss(c0 - e)) / (2 * h)
        print(f"    pixel {i}: adjoint {gf[i]:.6e} fd {fd:.6e}", flush=True)
        assert abs(fd - gf[i]) <= 1e-3 * np.abs(gf).max()


def test_grad_c0_rejects_unsupported():
    kg, med, src, sen, kw, obs = _setup("hetc_absorb_psrc", (32, 32), 10)
    m = copy.deepcopy(med)
    m.BonA = 5.0
    with pytest.raises(K.KWaveHIPError):
        K.gradient_c0(copy.deepcopy(kg), m, src, sen, obs, data_cast=np.float64, **kw)
    with pytest.raises(ValueError):
        K.gradient_c0(copy.deepcopy(kg), med, src, sen, obs[:, :5], data_cast=np.float64, **kw)
