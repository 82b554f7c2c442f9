"""The fused y-inverse / complex-to-real plane pass (k_plane_c2r) against the oracle.

3-D FP32 grids with 256 x 256 planes run the y inverse transform and the
complex-to-real x transform of the equation-of-state term in one pass per z
plane (the DC and Nyquist columns share one complex y transform).  The checks
are those of the unfused path: one step from a random state for each EOS
branch (FP32 to 1e-5 relative L2 per field, as in test_gpu_step), and a
40-step heterogeneous absorbing nonlinear run against the oracle (1e-4).
"""
import copy

import numpy as np
import pytest

import kwave_amd as K
from oracle import kwave_oracle as O

from cases import rel_l2, scenario
from test_gpu_step import _problem

pytestmark = pytest.mark.gpu

SHAPE = (256, 256, 4)
STEP_CASES = {
    "het_absorb_nonlinear": dict(shape=SHAPE, het=True, alpha="stokes", bona=True, src=True),
    "hom_stokes_nonlinear": dict(shape=SHAPE, het=False, alpha="stokes", bona=True, src=True),
    "hom_lossless": dict(shape=SHAPE, het=False, alpha=None, bona=False, src=False),
}


def test_plane_pass_in_use():
    """FP32 256 x 256 planes take the fused pass; FP64 and other planes the separate ones."""
    kg, med, src, sen, _ = _problem(STEP_CASES["hom_lossless"])
    for dtype, fused in ((np.float32, True), (np.float64, False)):
        desc, keep, _ = K.build_desc(copy.deepcopy(kg), med, src, sen, data_cast=dtype, pml_size=4)
        with K.Simulation(desc, keep) as sim:
            names = set(sim.profile(1))
        print(f"\n[case] {dtype.__name__}: {sorted(names)}", flush=True)
        assert ("k_plane_c2r" in names) == fused
        assert ("k_c2r_x" in names) == (not fused)


@pytest.mark.parametrize("name", list(STEP_CASES))
def test_plane_one_step_from_state(name):
    print(f"\n[case] plane pass, one step from a given state: {name} {SHAPE}", flush=True)
    c = STEP_CASES[name]
    kg, med, src, sen, state = _problem(c)
    kw = dict(pml_size=4, pml_alpha=2.0, smooth_p0=False)
    ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, n_steps=1, initial_state=state, **kw)["_state"]
    desc, keep, _ = K.build_desc(copy.deepcopy(kg), med, src, sen, data_cast=np.float32, **kw)
    stokes = c["alpha"] == "stokes" and not c["het"]
    with K.Simulation(desc, keep) as sim:
        sim.set_field("p", state["p"])
        for a in range(3):
            sim.set_field("u" + "xyz"[a], state["u"][a])
            sim.set_field("rho" + "xyz"[a], state["rho"][a])
        if stokes:
            sim.set_field("rho_prev", state["rho_prev"])
        sim.run(1)
        got = {k: sim.field(k) for k in ("p", "ux", "uy", "uz", "rhox", "rhoy", "rhoz")}
    want = {"p": ref["p"]}
    for a in range(3):
        want["u" + "xyz"[a]] = ref["u"][a]
        want["rho" + "xyz"[a]] = ref["rho"][a]
    for k in want:
        e = rel_l2(got[k], np.asarray(want[k], dtype=np.float64))
        print(f"    {k:5s} rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-5, (k, e)


def test_plane_parity_run():
    name, shape = "hetc_absorb_nonlin_psrc", (256, 256, 6)
    print(f"\n[case] plane pass, {name} {shape} float32, 40 steps", flush=True)
    kg, med, src, sen, kw = scenario(name, shape, nt=40)
    ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
    out = K.kspace_first_order(kg, med, src, sen, data_cast=np.float32, **kw)
    for k in out.keys():
        if k.startswith("_"):
            continue
        e = rel_l2(out[k], ref[k])
        print(f"    {k:8s} rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-4, (k, e)
