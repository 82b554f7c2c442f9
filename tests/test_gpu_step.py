"""One time step from a given state (kw_sim_set_field + kw_sim_run(sim, 1)) against the oracle.

The operator-level form of time_step_{1,2,3}d! (first_order_steps.jl:296-535):
a random state (p, u_d, rho_d and, for the homogeneous :stokes branch, the
previous total density) is written into the simulation, one step is run, and
every field must equal one oracle step from the same state -- FP64 kernels to
1e-12 and FP32 kernels to 1e-5 relative L2 per field -- for each EOS branch,
with PML, heterogeneous and scalar maps, B/A, and sources active at step 1.
"""
import copy

import numpy as np
import pytest

import kwave_amd as K
from oracle import kwave_oracle as O

from cases import rel_l2

pytestmark = pytest.mark.gpu

CASES = {
    "het_absorb_nonlinear_3d": dict(shape=(24, 20, 16), het=True, alpha="stokes", bona=True, src=True),
    "hom_stokes_nonlinear_2d": dict(shape=(40, 32), het=False, alpha="stokes", bona=True, src=True),
    "hom_nodisp_2d": dict(shape=(36, 32), het=False, alpha="no_dispersion", bona=False, src=False),
    "het_lossless_1d": dict(shape=(64,), het=True, alpha=None, bona=False, src=True),
    "hom_lossless_3d": dict(shape=(16, 24, 20), het=False, alpha=None, bona=False, src=False),
}


def _problem(c):
    shape = c["shape"]
    nd = len(shape)
    rng = np.random.default_rng(sum(shape))
    args = []
    for n in shape:
        args += [n, 1e-4]
    kg = K.KWaveGrid(*args)
    if c["het"]:
        c0 = 1500.0 + 60.0 * rng.random(shape)
        rho0 = 1000.0 + 100.0 * rng.random(shape)
    else:
        c0, rho0 = 1520.0, 1010.0
    kw = dict(sound_speed=c0, density=rho0)
    if c["alpha"]:
        kw.update(alpha_coeff=2.0, alpha_power=1.4, alpha_mode=c["alpha"])
    if c["bona"]:
        kw.update(BonA=6.0 if not c["het"] else 4.0 + rng.random(shape))
    med = K.KWaveMedium(**kw)
    K.make_time(kg, 1700.0)
    kg.Nt = 3
    kg.t_array = np.arange(3) * kg.dt
    src = K.KWaveSource()
    if c["src"]:
        m = np.zeros(int(np.prod(shape)), dtype=bool)
        m[rng.choice(m.size, size=6, replace=False)] = True
        m = m.reshape(shape, order="F")
        src = K.KWaveSource(p_mask=m, p=1e4 * rng.standard_normal((6, 3)))
    sen = K.KWaveSensor(mask=np.ones(shape, dtype=bool), record=["p"])
    state = {"p": 1e5 * rng.standard_normal(shape),
             "u": [0.05 * rng.standard_normal(shape) for _ in range(nd)],
             "rho": [0.02 * rng.standard_normal(shape) for _ in range(nd)],
             "rho_prev": 0.02 * nd * rng.standard_normal(shape)}
    return kg, med, src, sen, state


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("name", list(CASES))
def test_one_step_from_state(name, dtype):
    print(f"\n[case] one step from a given state: {name}, {dtype.__name__}", flush=True)
    c = CASES[name]
    kg, med, src, sen, state = _problem(c)
    nd = len(c["shape"])
    kw = dict(pml_size=4, pml_alpha=2.0, smooth_p0=False)
    ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, n_steps=1, initial_state=state, **kw)["_state"]
    desc, keep, _ = K.build_desc(copy.deepcopy(kg), med, src, sen, data_cast=dtype, **kw)
    stokes = c["alpha"] == "stokes" and not c["het"]
    with K.Simulation(desc, keep) as sim:
        sim.set_field("p", state["p"])
        for a in range(nd):
            sim.set_field("u" + "xyz"[a], state["u"][a])
            sim.set_field("rho" + "xyz"[a], state["rho"][a])
        if stokes:
            sim.set_field("rho_prev", state["rho_prev"])
        sim.run(1)
        got = {"p": sim.field("p")}
        for a in range(nd):
            got["u" + "xyz"[a]] = sim.field("u" + "xyz"[a])
            got["rho" + "xyz"[a]] = sim.field("rho" + "xyz"[a])
    want = {"p": ref["p"]}
    for a in range(nd):
        want["u" + "xyz"[a]] = ref["u"][a]
        want["rho" + "xyz"[a]] = ref["rho"][a]
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for k in want:
        e = rel_l2(got[k], np.asarray(want[k], dtype=np.float64))
        print(f"    {k:5s} rel-L2 {e:.3e}", flush=True)
        assert e <= tol, (k, e)


def test_set_field_rejects_absent_fields():
    kg, med, src, sen, state = _problem(CASES["het_lossless_1d"])
    desc, keep, _ = K.build_desc(copy.deepcopy(kg), med, src, sen, pml_size=4)
    with K.Simulation(desc, keep) as sim:
        with pytest.raises(K.KWaveHIPError):
            sim.set_field("uy", state["p"])          # 1-D grid has no uy
        with pytest.raises(K.KWaveHIPError):
            sim.set_field("rho_prev", state["p"])    # lossless: no previous density
        with pytest.raises(ValueError):
            sim.set_field("p", state["p"][:10])
