"""Pins of the oracle's discrete adjoint (dL/dc0) on the CPU.

The adjoint is checked against what fixes it independently of its own
formulas: central finite differences of the oracle's forward solver (the
restatement of first_order.jl that the golden vectors pin), and the loss it
reports against the forward solver's sensor data.  A planted mistake (the
c_ref = mean(c0) coupling term dropped) must fail the finite-difference check.
"""
import copy

import numpy as np
import pytest

from oracle import kwave_oracle as O

from cases import scenario


def _setup(name, shape, nt, seed=3):
    kg, med, src, sen, kw = scenario(name, shape, seed=seed, nt=nt)
    sen.record = ["p"]
    rng = np.random.default_rng(seed + 100)
    c_true = np.asarray(med.sound_speed, dtype=np.float64) + 30.0 * rng.standard_normal(shape)
    med_true = copy.deepcopy(med)
    med_true.sound_speed = c_true
    obs = O.kspace_first_order(copy.deepcopy(kg), med_true, src, sen, **kw)["p"]
    return kg, med, src, sen, kw, obs


def _loss(kg, med, src, sen, kw, obs, c):
    m = copy.deepcopy(med)
    m.sound_speed = c
    p = O.kspace_first_order(copy.deepcopy(kg), m, src, sen, **kw)["p"]
    return 0.5 * float(np.sum((p - obs) ** 2))


CASES = [
    ("hetc_absorb_psrc", (24, 20), 30),
    ("hetc_hetr_psrc_p0", (20, 24), 25),
    ("hetc_absorb_psrc", (40,), 40),
    ("hetc_hetr_psrc", (10, 12, 8), 16),
]


@pytest.mark.parametrize("name,shape,nt", CASES, ids=[f"{c[0]}-{len(c[1])}d" for c in CASES])
def test_adjoint_matches_finite_differences(name, shape, nt):
    kg, med, src, sen, kw, obs = _setup(name, shape, nt)
    loss, g = O.grad_c0(copy.deepcopy(kg), med, src, sen, obs, pml_size=kw["pml_size"],
                        pml_alpha=kw["pml_alpha"], smooth_p0=kw["smooth_p0"])
    c0 = np.asarray(med.sound_speed, dtype=np.float64)
    assert loss == pytest.approx(_loss(kg, med, src, sen, kw, obs, c0), rel=1e-12)
    rng = np.random.default_rng(5)
    flat = np.argsort(-np.abs(g).ravel(order="F"))[:2].tolist() + rng.choice(c0.size, 2, replace=False).tolist()
    gmax = np.abs(g).max()
    for i in flat:
        h = 1e-5 * c0.ravel(order="F")[i]
        e = np.zeros(c0.size)
        e[i] = h
        e = e.reshape(c0.shape, order="F")
        fd = (_loss(kg, med, src, sen, kw, obs, c0 + e) - _loss(kg, med, src, sen, kw, obs, c0 - e)) / (2 * h)
        assert abs(fd - g.ravel(order="F")[i]) <= 1e-6 * gmax, (i, fd, g.ravel(order="F")[i])


def test_adjoint_planted_mistake_fails(monkeypatch):
    """Dropping dK/dc_ref (the 1/N term through c_ref = mean(c0)) must break the FD match."""
    kg, med, src, sen, kw, obs = _setup("hetc_absorb_psrc", (24, 20), 30)
    c0 = np.asarray(med.sound_speed, dtype=np.float64)
    orig = O._eos_multiplier

    def no_cref(g, medium, c_ref, dt, derivative=False):
        m = orig(g, medium, c_ref, dt, derivative)
        return 0 * m if derivative else m

    monkeypatch.setattr(O, "_eos_multiplier", no_cref)
    _, g = O.grad_c0(copy.deepcopy(kg), med, src, sen, obs, pml_size=kw["pml_size"], pml_alpha=kw["pml_alpha"],
                     smooth_p0=False)
    monkeypatch.setattr(O, "_eos_multiplier", orig)
    i = int(np.random.default_rng(1).integers(c0.size))
    h = 1e-5 * c0.ravel(order="F")[i]
    e = np.zeros(c0.size)
    e[i] = h
    e = e.reshape(c0.shape, order="F")
    fd = (_loss(kg, med, src, sen, kw, obs, c0 + e) - _loss(kg, med, src, sen, kw, obs, c0 - e)) / (2 * h)
    # the dropped term shifts every pixel by ~4e-5 max|g|, 40x the FD check's tolerance
    assert abs(fd - g.ravel(order="F")[i]) > 1e-5 * np.abs(g).max()
