"""GPU parity with the line-FFT plan family forced through the tuning table.

Line lengths 256 and 512 have three compiled plan families (radix schedule
and threads per line); the tuning table's optional fifth column picks one per
(precision, pass, length).  Each family on every pass must give the oracle's
fields to round-off in FP64, forward and adjoint, and the autotuner's choice
must keep parity too.
"""
import copy

import numpy as np
import pytest

import kwave_amd as K
from kwave_amd import _lib, tuning
from oracle import kwave_oracle as O

from cases import rel_l2, scenario

pytestmark = pytest.mark.gpu

PASSES = range(6)   # dir_x, dir_y, dir_z, spec_y, spec_z, c2r_x


def _force(plan, lengths, lpw=4, prec=(4, 8)):
    text = "".join(f"{p} {c} {n} {lpw} {plan}\n" for p in prec for c in PASSES for n in lengths)
    tuning.clear_table()
    _lib.check(_lib.load().kw_tune_import(text.encode()))
    assert len(tuning.plans()) == len(prec) * len(PASSES) * len(lengths)


FWD = [("hetc_absorb_psrc", (256, 256), 24), ("hetc_hetr_nonlin_psrc_urec", (512, 256), 16),
       ("hetc_absorb_p0", (16, 256, 256), 10), ("stokes_p0_psrc", (512,), 60)]


@pytest.mark.parametrize("plan", [0, 1, 2])
@pytest.mark.parametrize("name,shape,nt", FWD, ids=[f"{len(c[1])}d-{'x'.join(map(str, c[1]))}" for c in FWD])
def test_forced_plan_family_fp64(plan, name, shape, nt):
    print(f"\n[case] plan family {plan} fp64 {name} {shape} nt={nt}", flush=True)
    kg, med, src, sen, kw = scenario(name, shape, seed=5, nt=nt)
    try:
        _force(plan, (256, 512))
        out = K.kspace_first_order(copy.deepcopy(kg), med, src, sen, data_cast=np.float64, **kw)
    finally:
        tuning.clear_table()
    ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
    for k in out.keys():
        e = rel_l2(out[k], ref[k])
        print(f"    {k:8s} rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-10, (k, e)


@pytest.mark.parametrize("plan", [0, 1, 2])
def test_forced_plan_family_fp32_3d(plan):
    """FP32 3-D: the fused plane pass replaces the y-inverse and c2r passes where it applies."""
    print(f"\n[case] plan family {plan} fp32 3d 16x256x256", flush=True)
    kg, med, src, sen, kw = scenario("hetc_absorb_psrc", (16, 256, 256), seed=6, nt=10)
    try:
        _force(plan, (256,))
        out = K.kspace_first_order(copy.deepcopy(kg), med, src, sen, data_cast=np.float32, **kw)
    finally:
        tuning.clear_table()
    ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
    for k in out.keys():
        e = rel_l2(out[k], ref[k])
        print(f"    {k:8s} rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-4, (k, e)


@pytest.mark.parametrize("plan", [0, 1, 2])
def test_forced_plan_family_adjoint_fp64(plan):
    print(f"\n[case] plan family {plan} adjoint fp64 2d 256x256", flush=True)
    kg, med, src, sen, kw = scenario("hetc_absorb_psrc", (256, 256), seed=3, nt=30)
    sen.record = ["p"]
    rng = np.random.default_rng(103)
    med_true = copy.deepcopy(med)
    med_true.sound_speed = np.asarray(med.sound_speed, dtype=np.float64) + 30.0 * rng.standard_normal((256, 256))
    obs = O.kspace_first_order(copy.deepcopy(kg), med_true, src, sen, **kw)["p"]
    loss_ref, g_ref = O.grad_c0(copy.deepcopy(kg), med, src, sen, obs, pml_size=kw["pml_size"],
                                pml_alpha=kw["pml_alpha"], smooth_p0=kw["smooth_p0"])
    try:
        _force(plan, (256,))
        loss, g = K.gradient_c0(copy.deepcopy(kg), med, src, sen, obs, checkpoint_every=8, data_cast=np.float64,
                                **kw)
    finally:
        tuning.clear_table()
    e = rel_l2(g, g_ref)
    print(f"    loss {loss:.6e} ref {loss_ref:.6e}; grad rel-L2 {e:.3e}", flush=True)
    assert loss == pytest.approx(loss_ref, rel=1e-10)
    assert e <= 1e-9


def test_autotune_plan_trials_then_parity():
    """At length 256 the autotuner also times the plan families; whatever it keeps, parity holds."""
    print("\n[case] autotune 2d 256x256 f64 with plan trials, then parity", flush=True)
    kg, med, src, sen, kw = scenario("hetc_absorb_psrc", (256, 256), seed=8, nt=20)
    tuning.clear_table()
    try:
        t = tuning.autotune(copy.deepcopy(kg), med, src, sen, steps=3, data_cast=np.float64, **kw)
        pl = tuning.plans()
        print(f"    table {t}\n    plans {pl}", flush=True)
        assert {(8, "dir_x", 256), (8, "dir_y", 256), (8, "spec_y", 256), (8, "c2r_x", 256)} <= set(t)
        assert set(pl) <= set(t) and all(v in (0, 1, 2) for v in pl.values())
        out = K.kspace_first_order(copy.deepcopy(kg), med, src, sen, data_cast=np.float64, **kw)
    finally:
        tuning.clear_table()
    ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
    for k in out.keys():
        assert rel_l2(out[k], ref[k]) <= 1e-10, k


@pytest.mark.parametrize("plan", [0, 1, 2])
def test_forced_plan_family_axisymmetric_fp64(plan):
    """The axisymmetric solver's radial pass (1/r term) under each plan family."""
    print(f"\n[case] plan family {plan} axisymmetric fp64 256x256", flush=True)
    kg, med, src, sen, kw = scenario("hetc_hetr_absorb_psrc", (256, 256), seed=21, nt=24)
    try:
        _force(plan, (256,))
        out = K.kspace_first_order_as(copy.deepcopy(kg), med, src, sen, data_cast=np.float64, **kw)
    finally:
        tuning.clear_table()
    ref = O.kspace_first_order_as(copy.deepcopy(kg), med, src, sen, **kw)
    for k in out.keys():
        e = rel_l2(out[k], ref[k])
        print(f"    {k:8s} rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-10, (k, e)


@pytest.mark.parametrize("plan", [0, 1, 2])
def test_forced_plan_family_stacked_firings_fp64(plan):
    """Stacked firings (one launch per pass over all firings) under each plan family equal
    the default-family single runs."""
    print(f"\n[case] plan family {plan} stacked firings fp64 256x256", flush=True)
    kg = K.KWaveGrid(256, 1e-4, 256, 1e-4)
    med = K.KWaveMedium(sound_speed=1500.0 + 40.0 * np.random.default_rng(3).standard_normal((256, 256)),
                        density=1000.0, alpha_coeff=0.75, alpha_power=1.5, alpha_mode="stokes")
    K.make_time(kg, 1700.0)
    kg.Nt = 40
    kg.t_array = np.arange(kg.Nt) * kg.dt
    srcs, sens = [], []
    for i in range(3):
        m = np.zeros((256, 256), dtype=bool)
        m[100 + 3 * i, 60 + 30 * i] = True
        sig = 1e5 * K.tone_burst(1 / kg.dt, 1.5e6, 2)
        srcs.append(K.KWaveSource(p_mask=m, p=sig[None, :]))
        sm = np.zeros((256, 256), dtype=bool)
        sm[110, 20:230 - 10 * i] = True   # within the distance the wave covers in 40 steps
        sens.append(K.KWaveSensor(mask=sm, record=["p", "p_max"]))
    kw = dict(pml_size=12, pml_alpha=2.0, data_cast=np.float64)
    refs = [K.kspace_first_order(copy.deepcopy(kg), med, s, r, smooth_p0=False, **kw) for s, r in zip(srcs, sens)]
    try:
        _force(plan, (256,))
        outs = K.kspace_first_order_firings(copy.deepcopy(kg), med, srcs, sens, **kw)
    finally:
        tuning.clear_table()
    for i in range(3):
        for f in ("p", "p_max"):
            e = rel_l2(outs[i][f], refs[i][f])
            print(f"    firing {i} {f:6s} rel-L2 {e:.3e}", flush=True)
            assert e <= 1e-10, (i, f, e)
