"""Host-side checks of the C ABI and the Python mirror (no GPU compute)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import kwave_amd as K
from kwave_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kwave_hip.h")


def _declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(kw_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_the_binding_table():
    assert set(_declared_functions()) == set(_lib.EXPORTS)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libkwave_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    for name in _declared_functions():
        assert hasattr(lib, name), name
    lib = _lib.load()
    assert lib.kw_abi_version() == 1
    for n in (8, 9, 15, 24, 25, 27, 45, 81, 100, 125, 128, 135, 225, 243, 256, 384, 400, 512, 768, 1024, 4096):
        assert lib.kw_fft_length_supported(n) == 1
    for n in (2, 7, 11, 101, 257, 1000, 2048 - 1):
        assert lib.kw_fft_length_supported(n) == 2     # chirp-z lines on a power-of-two plan
    for n in (0, 2049, 3001, 8192):
        assert lib.kw_fft_length_supported(n) == 0


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libkwave_hip.so not built")
def test_grid_lane_offset_limit():
    """Strided lines spanning more than 4 GiB are rejected (32-bit lane offsets)."""
    lib = _lib.load()

    def ok(shape, prec):
        n = (C.c_int64 * 3)(*(list(shape) + [1] * (3 - len(shape))))
        return lib.kw_grid_supported(len(shape), n, prec)

    assert ok((256, 256, 256), 4) == 1 and ok((256, 256, 256), 8) == 1
    assert ok((512, 512, 512), 4) == 1 and ok((512, 512, 512), 8) == 1
    assert ok((1024, 1024, 1024), 4) == 0          # spectral z line: 528*1024*1023 complex = 4.4e9 B
    assert ok((1024, 1024, 512), 8) == 0           # 528*1024*511 complex128 = 4.4e9 B
    assert ok((2048, 2048, 256), 4) == 0
    assert ok((4096, 4096), 8) == 1                # 2-D lines stay far below the limit
    assert ok((1000, 1000), 4) == 1                # chirp-z lines
    assert ok((3001, 64), 4) == 0                  # longer than the chirp-z limit without a plan
    assert ok((256, 256, 256), 2) == 0             # unknown precision


def test_ctypes_descriptor_layout_matches_header():
    """Compile the header with gcc and compare sizeof/offsetof with ctypes."""
    fields = [f[0] for f in _lib.SimDesc._fields_]
    src = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){",
           'printf("size %zu\\n", sizeof(kw_sim_desc));']
    for f in fields:
        src.append(f'printf("{f} %zu\\n", offsetof(kw_sim_desc, {f}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c99", c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    vals = dict(l.split() for l in out if l)
    assert int(vals["size"]) == C.sizeof(_lib.SimDesc)
    for f in fields:
        assert int(vals[f]) == getattr(_lib.SimDesc, f).offset, f


def _small():
    kg = K.KWaveGrid(32, 1e-4, 32, 1e-4)
    med = K.KWaveMedium(sound_speed=1500.0, density=1000.0)
    src = K.KWaveSource(p0=np.zeros((32, 32)))
    m = np.zeros((32, 32), bool)
    m[3, 4] = True
    return kg, med, src, m


def test_validation_errors_match_reference():
    kg, med, src, m = _small()
    with pytest.raises(ValueError, match="Time array not set"):                   # first_order.jl:311-313
        K.kspace_first_order(kg, med, src, K.KWaveSensor(mask=m))
    K.make_time(kg, 1500.0)
    with pytest.raises(ValueError, match="Invalid sensor record field"):          # sensor.jl:54-56
        K.kspace_first_order(kg, med, src, K.KWaveSensor(mask=m, record=["p", "pressure"]))
    with pytest.raises(TypeError):                                                # first_order.jl:294
        K.build_desc(kg, med, src, K.KWaveSensor(mask=m), pml_alpha=2)
    with pytest.raises(TypeError):
        K.build_desc(kg, med, src, K.KWaveSensor(mask=m), data_cast=np.float16)


def test_descriptor_contents():
    kg, med, src, m = _small()
    K.make_time(kg, 1500.0)
    c0 = np.full((32, 32), 1500.0)
    c0[5:, :] = 1600.0
    med = K.KWaveMedium(sound_speed=c0, density=1000.0, alpha_coeff=0.75, alpha_power=1.5, alpha_mode="stokes",
                        BonA=5.0)
    pm = np.zeros((32, 32), bool)
    pm[4, 7] = pm[9, 2] = True
    src = K.KWaveSource(p_mask=pm, p=np.ones((2, 10)))
    d, keep, info = K.build_desc(kg, med, src, K.KWaveSensor(mask=m, record=["p", "p_rms"]),
                                 pml_size=(10, 12), pml_alpha=(2.0, 3.0), data_cast=np.float32)
    assert d.ndim == 2 and tuple(d.N) == (32, 32, 1) and d.precision == 4
    assert tuple(d.pml_size)[:2] == (10, 12) and tuple(d.pml_alpha)[:2] == (2.0, 3.0)
    assert d.c_ref == pytest.approx(np.mean(c0)) and d.absorbing == 1 and d.alpha_mode == 2
    assert d.nonlinear == 1 and d.bona_scalar == 5.0
    assert d.n_psrc == 2 and [d.psrc_index[i] for i in range(2)] == [9 + 32 * 2, 4 + 32 * 7]  # findall order
    assert d.record == _lib.REC_BITS["p"] | _lib.REC_BITS["p_rms"] and d.n_sensor == 1
    assert d.sensor_index[0] == 3 + 32 * 4


def test_product_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    kg, med, src, m = _small()
    K.make_time(kg, 1500.0)
    with pytest.raises(_lib.KWaveHIPError, match="missing"):
        K.kspace_first_order(kg, med, src, K.KWaveSensor(mask=m), smooth_p0=False)


def test_simulation_output_mirror():
    out = K.SimulationOutput({"p": np.zeros((2, 3)), "p_max": np.zeros(2)})
    assert out[":p"].shape == (2, 3) and "p_max" in out and out.haskey("p") and set(out.keys()) == {"p", "p_max"}
    assert repr(out) == "SimulationOutput with fields: p, p_max"


def test_tuning_table_text_round_trip():
    """kw_tune_import / kw_tune_export / kw_tune_clear need no device."""
    from kwave_amd import _lib, tuning
    lib = _lib.load()
    tuning.clear_table()
    try:
        _lib.check(lib.kw_tune_import(b"4 0 256 8\n8 3 96 2\n"))
        assert tuning.table() == {(4, "dir_x", 256): 8, (8, "spec_y", 96): 2}
        with pytest.raises(_lib.KWaveHIPError):
            _lib.check(lib.kw_tune_import(b"4 0 257 8\n"))       # no plan for 257
        with pytest.raises(_lib.KWaveHIPError):
            _lib.check(lib.kw_tune_import(b"4 9 256 8\n"))       # no such pass
        with pytest.raises(_lib.KWaveHIPError):
            _lib.check(lib.kw_tune_import(b"4 0 256 x\n"))
        assert len(tuning.table()) == 2                          # failed imports add nothing
        assert tuning.plans() == {}
        # optional fifth column: plan family, only where the line length has one
        _lib.check(lib.kw_tune_import(b"4 5 512 4 0\n\n8 1 256 16 1\n"))
        assert tuning.plans() == {(4, "c2r_x", 512): 0, (8, "dir_y", 256): 1}
        assert tuning.table()[(4, "c2r_x", 512)] == 4
        assert "4 5 512 4 0" in tuning.table_text().splitlines()
        for bad in (b"4 0 96 8 1\n",        # no alternative plans at 96
                    b"4 0 256 8 3\n",       # no plan family 3
                    b"4 0 256 8 1 7\n",     # six columns
                    b"4 0 256\n"):          # three columns
            with pytest.raises(_lib.KWaveHIPError):
                _lib.check(lib.kw_tune_import(bad))
        assert len(tuning.plans()) == 2
        _lib.check(lib.kw_tune_import(b"4 5 512 4\n"))          # re-import without a plan: the default again
        assert tuning.plans() == {(8, "dir_y", 256): 1}
    finally:
        tuning.clear_table()
    assert tuning.table() == {} and tuning.plans() == {}


def test_errors_from_every_translation_unit_reach_kw_last_error():
    """The operator and elastic entry points report through the same thread-local message."""
    import kwave_amd as K
    from kwave_amd import operators
    with pytest.raises(K.KWaveHIPError, match="unknown axis table"):
        lib = _lib.load()
        out = (C.c_double * 8)()
        _lib.check(lib.kw_test_axis_table(8, 1e-4, 0, 0.0, 1e-8, 99, out))
    with pytest.raises(K.KWaveHIPError, match="pml_size out of range"):
        operators.axis_table(8, 1e-4, "pml", pml_size=9, pml_alpha=2.0, dt=1e-8)
    lib = _lib.load()
    assert lib.kw_elastic_create(None, None, None) == _lib.KW_ERR_INVALID
    assert b"null argument" in lib.kw_last_error()


def test_time_reversal_data_rows_checked():
    """Fewer TR-data rows than sensor points is the reference's BoundsError (first_order.jl:455)."""
    kg, med, _, m = _small()
    K.make_time(kg, 1500.0)
    m = m.copy()
    m[10, 10] = True                                     # two sensor points
    sen = K.KWaveSensor(mask=m, time_reversal_boundary_data=np.zeros((1, 40)))
    with pytest.raises(IndexError, match="fewer rows"):
        K.build_desc(kg, med, K.KWaveSource(), sen)
    sen = K.KWaveSensor(mask=m, time_reversal_boundary_data=np.zeros((2, 40)))
    d, keep, info = K.build_desc(kg, med, K.KWaveSource(), sen)
    assert info["tr"] and d.Nt == 40 and d.n_sensor == 2


def test_tuning_table_file_round_trip_keeps_plan_families(tmp_path):
    """save_table / load_table carry the optional plan-family column unchanged."""
    from kwave_amd import _lib, tuning
    lib = _lib.load()
    tuning.clear_table()
    try:
        _lib.check(lib.kw_tune_import(b"4 0 256 2 2\n4 3 512 8\n8 5 256 4 0\n"))
        before_t, before_p = tuning.table(), tuning.plans()
        path = str(tmp_path / "kw_tune.txt")
        tuning.save_table(path)
        tuning.clear_table()
        assert tuning.table() == {} and tuning.plans() == {}
        tuning.load_table(path)
        assert tuning.table() == before_t and tuning.plans() == before_p
        assert before_p == {(4, "dir_x", 256): 2, (8, "c2r_x", 256): 0}
    finally:
        tuning.clear_table()
