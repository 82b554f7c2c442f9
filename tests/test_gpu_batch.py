"""GPU parity of the batched form (many simulations per launch) vs the oracle."""
import copy

import numpy as np
import pytest

import kwave_amd as K
from oracle import kwave_oracle as O

from cases import rel_l2, scenario

pytestmark = pytest.mark.gpu


def _sims(names_shapes, nt):
    out = []
    for i, (name, shape) in enumerate(names_shapes):
        kg, med, src, sen, kw = scenario(name, shape, seed=11 + i, nt=nt)
        out.append((kg, med, src, sen))
    return out, kw


@pytest.mark.parametrize("use_graph", [True, False], ids=["graph", "stream"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_batch_2d_uneven(use_graph, dtype):
    print(f"\n[case] batch 2d graph={use_graph} {dtype.__name__}", flush=True)
    sims, kw = _sims([("hetc_absorb_psrc", (64, 64)), ("hetc_hetr_nonlin_psrc_urec", (96, 96)),
                      ("psrc", (128, 64)), ("stokes_p0_psrc", (64, 96)), ("hetc_absorb_psrc", (64, 64)),
                      ("usrc_urec", (48, 48))], nt=40)
    outs = K.kspace_first_order_batch([tuple(copy.deepcopy(s)) for s in sims], data_cast=dtype,
                                      use_graph=use_graph, **kw)
    tol = 1e-10 if dtype == np.float64 else 1e-4
    for i, (kg, med, src, sen) in enumerate(sims):
        ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
        for k in outs[i].keys():
            e = rel_l2(outs[i][k], ref[k])
            print(f"    sim {i} {k:8s} rel-L2 {e:.3e}", flush=True)
            assert e <= tol, (i, k, e)


def test_batch_3d():
    print("\n[case] batch 3d", flush=True)
    sims, kw = _sims([("hetc_absorb_nonlin_psrc", (32, 24, 16)), ("psrc", (16, 32, 24)),
                      ("hetc_hetr_usrc_urec_p0", (32, 24, 16))], nt=25)
    outs = K.kspace_first_order_batch([tuple(copy.deepcopy(s)) for s in sims], data_cast=np.float64, **kw)
    for i, (kg, med, src, sen) in enumerate(sims):
        ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
        for k in outs[i].keys():
            e = rel_l2(outs[i][k], ref[k])
            assert e <= 1e-10, (i, k, e)


def _table_for(lengths, lpw):
    return "".join(f"{p} {c} {n} {lpw}\n" for p in (4, 8) for c in range(6) for n in lengths)


def test_batch_with_tuned_shapes():
    """Non-default lines per workgroup (through the tuning table) keep parity."""
    print("\n[case] batch 2d f64 with 2 lines per workgroup", flush=True)
    from kwave_amd import tuning
    sims, kw = _sims([("hetc_absorb_psrc", (64, 64)), ("stokes_p0_psrc", (64, 96)), ("psrc", (128, 48))], nt=30)
    tuning.clear_table()
    try:
        from kwave_amd import _lib
        _lib.check(_lib.load().kw_tune_import(_table_for((48, 64, 96, 128), 2).encode()))
        outs = K.kspace_first_order_batch([tuple(copy.deepcopy(s)) for s in sims], data_cast=np.float64, **kw)
        single = K.kspace_first_order(*copy.deepcopy(sims[0]), data_cast=np.float64, **kw)
    finally:
        tuning.clear_table()
    for i, (kg, med, src, sen) in enumerate(sims):
        ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
        for k in outs[i].keys():
            assert rel_l2(outs[i][k], ref[k]) <= 1e-10, (i, k)
    assert rel_l2(single["p"], O.kspace_first_order(copy.deepcopy(sims[0][0]), *sims[0][1:], **kw)["p"]) <= 1e-10


def test_autotune_fills_table():
    print("\n[case] autotune 2d 64x96 f32", flush=True)
    from kwave_amd import tuning
    tuning.clear_table()
    try:
        sims, kw = _sims([("hetc_absorb_psrc", (64, 96))], nt=20)
        kg, med, src, sen = sims[0]
        t = tuning.autotune(kg, med, src, sen, steps=3, data_cast=np.float32, **kw)
    finally:
        tuning.clear_table()
    print(f"    table {t}", flush=True)
    for key in [(4, "dir_x", 64), (4, "dir_y", 96), (4, "spec_y", 96), (4, "c2r_x", 64)]:
        assert key in t and t[key] >= 1


def test_autotune_batch_then_parity():
    print("\n[case] batch autotune (whole-batch trials) then f64 parity", flush=True)
    from kwave_amd import tuning
    sims, kw = _sims([("hetc_absorb_psrc", (64, 64)), ("stokes_p0_psrc", (64, 96)), ("psrc", (128, 48))], nt=24)
    tuning.clear_table()
    try:
        t = tuning.autotune_batch([tuple(copy.deepcopy(s)) for s in sims], steps=3, data_cast=np.float64, **kw)
        print(f"    table {t}", flush=True)
        assert (8, "dir_y", 64) in t and (8, "spec_y", 96) in t
        outs = K.kspace_first_order_batch([tuple(copy.deepcopy(s)) for s in sims], data_cast=np.float64, **kw)
    finally:
        tuning.clear_table()
    for i, (kg, med, src, sen) in enumerate(sims):
        ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
        for k in outs[i].keys():
            assert rel_l2(outs[i][k], ref[k]) <= 1e-10, (i, k)


def test_batch_odd_and_long_lines():
    """Odd line lengths (unpaired strided loads) and 768/1024-point lines (XCD-ordered
    strided and spectral passes) in one batch, FP32 and FP64."""
    sims, kw = _sims([("hetc_absorb_psrc", (45, 27)), ("stokes_p0_psrc", (75, 25)),
                      ("hetc_hetr_nonlin_psrc_urec", (45, 27)), ("psrc", (16, 768)),
                      ("hetc_absorb_psrc", (24, 1024))], nt=30)
    for dtype, tol in ((np.float64, 1e-10), (np.float32, 1e-4)):
        print(f"\n[case] batch odd + long lines {dtype.__name__}", flush=True)
        outs = K.kspace_first_order_batch([tuple(copy.deepcopy(s)) for s in sims], data_cast=dtype,
                                          use_graph=True, **kw)
        for i, (kg, med, src, sen) in enumerate(sims):
            ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)
            for k in outs[i].keys():
                e = rel_l2(outs[i][k], ref[k])
                assert e <= tol, (i, k, e)


@pytest.mark.parametrize("grid8", [1, 0], ids=["grid8", "fill"])
def test_batch_persistent_grids(grid8):
    """Persistent-grid launches (each workgroup walking many tiles) keep parity; with an
    8-workgroup grid every batched pass loops over its tiles."""
    print(f"\n[case] batch persistent grid8={grid8} 2d f64 + 3d f64", flush=True)
    from kwave_amd.batch import Batch
    for names, nt in (([("hetc_absorb_psrc", (64, 64)), ("stokes_p0_psrc", (64, 96)), ("psrc", (128, 48)),
                        ("usrc_urec", (48, 48))], 30),
                      ([("hetc_absorb_nonlin_psrc", (32, 24, 16)), ("psrc", (16, 32, 24))], 20)):
        sims, kw = _sims(names, nt=nt)
        descs, keep = [], []
        for s in sims:
            d, k, _ = K.build_desc(*copy.deepcopy(s), data_cast=np.float64, **kw)
            descs.append(d)
            keep.append(k)
        for use_graph in (True, False):
            with Batch(descs, keep, use_graph=use_graph, persistent=True, grid8=grid8) as b:
                b.run(nt)
                b.synchronize()
                got = [b.sensor(i, "p") for i in range(len(sims))]
            for i, (kg, med, src, sen) in enumerate(sims):
                ref = O.kspace_first_order(copy.deepcopy(kg), med, src, sen, **kw)["p"]
                e = rel_l2(got[i], ref)
                print(f"    {len(names[0][1])}d graph={use_graph} sim {i} rel-L2 {e:.3e}", flush=True)
                assert e <= 1e-10, (i, use_graph, e)


def test_batch_flags_rejected():
    """Unknown kw_batch_create flag bits, and a grid cap without persistent grids, are KW_ERR_INVALID."""
    import ctypes as C
    from kwave_amd import _lib
    sims, kw = _sims([("psrc", (32, 32))], nt=4)
    d, k, _ = K.build_desc(*copy.deepcopy(sims[0]), data_cast=np.float64, **kw)
    arr = (_lib.SimDesc * 1)(d)
    lib = _lib.load()
    for flags in (4, 1 | (1 << 8), -1):
        h = C.c_void_p()
        with pytest.raises(_lib.KWaveHIPError):
            _lib.check(lib.kw_batch_create(_lib.context(0), arr, 1, flags, C.byref(h)))
    del k
