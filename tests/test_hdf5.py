"""HDF5 wire format (KWave.jl/src/io/hdf5.jl) and the CLI (KWave.jl/src/cli.jl), host side.

The reference holds no .h5 fixtures, so the writer's parity with HDF5.jl is
"parity unpinned"; what pins it here:
- round trips through the reader (dtype, Julia shape, values, attributes);
- byte-level checks against the HDF5 specification: superblock fields at their
  fixed offsets, B-tree/heap/symbol-node signatures at the addresses the
  superblock and object headers give, the chunk being one zlib stream of the
  column-major bytes;
- the IEEE float64 and string datatype encodings compared with the bytes a
  real libhdf5 wrote (scipy's MATLAB v7.3 test file, when installed), and the
  reader checked on that file (superblock at offset 512, symbol-table group,
  version-2 layout message, attribute).
"""
import os
import struct
import zlib

import numpy as np
import pytest

import kwave_amd as K
from kwave_amd import cli, hdf5

MAT73 = None
try:
    import scipy.io.matlab as _m

    _p = os.path.join(os.path.dirname(_m.__file__), "tests", "data", "testhdf5_7.4_GLNX86.mat")
    if os.path.exists(_p):
        MAT73 = _p
except ImportError:
    pass


def _u64(b, off):
    return struct.unpack_from("<Q", b, off)[0]


def test_round_trip_arrays_and_scalars(tmp_path):
    fn = str(tmp_path / "a.h5")
    rng = np.random.default_rng(1)
    a3 = rng.standard_normal((5, 4, 3))
    iv = np.arange(7, dtype=np.int64) - 3
    hdf5.write_matrix(fn, "a3", a3)
    hdf5.write_matrix(fn, "iv", iv, compression=False)
    hdf5.write_matrix(fn, "s", 2.5)
    hdf5.write_matrix(fn, "n", 42)
    hdf5.write_matrix(fn, "f32", a3.astype(np.float32)[:, :, 0])
    f = hdf5.read_file(fn)
    assert set(f.datasets) == {"a3", "iv", "s", "n", "f32"}
    got = hdf5.read_matrix(fn, "a3")
    assert got.dtype == np.float64 and got.shape == (5, 4, 3)
    np.testing.assert_array_equal(got, a3)
    np.testing.assert_array_equal(hdf5.read_matrix(fn, "iv"), iv)
    assert hdf5.read_matrix(fn, "iv").dtype == np.int64
    np.testing.assert_array_equal(hdf5.read_matrix(fn, "s"), [2.5])
    np.testing.assert_array_equal(hdf5.read_matrix(fn, "n"), [42])
    np.testing.assert_array_equal(hdf5.read_matrix(fn, "f32"), a3.astype(np.float32)[:, :, 0])
    assert f.datasets["a3"].attrs["data_type"] == "float"
    assert f.datasets["iv"].attrs["data_type"] == "long"
    assert f.datasets["n"].attrs["data_type"] == "long"
    # hdf5.jl:25: one deflate chunk only when compressing more than one element
    assert f.datasets["a3"].chunked and not f.datasets["iv"].chunked and not f.datasets["s"].chunked


def test_rewrite_replaces_dataset_and_keeps_others(tmp_path):
    fn = str(tmp_path / "b.h5")
    hdf5.write_matrix(fn, "x", np.ones(3))
    hdf5.write_matrix(fn, "y", np.zeros((2, 2)))
    hdf5.write_matrix(fn, "x", np.arange(5.0))
    f = hdf5.read_file(fn)
    np.testing.assert_array_equal(f.datasets["x"].value, np.arange(5.0))
    np.testing.assert_array_equal(f.datasets["y"].value, np.zeros((2, 2)))


def test_many_datasets_span_several_symbol_nodes(tmp_path):
    fn = str(tmp_path / "m.h5")
    flags = {f"flag_{i:02d}": i for i in range(20)}
    flags["name"] = "value"
    hdf5.write_flags(fn, flags)
    f = hdf5.read_file(fn)
    assert len(f.datasets) == 21
    for i in range(20):
        v = f.datasets[f"flag_{i:02d}"].value
        assert v.dtype == np.int64 and list(v) == [i]
    assert f.datasets["name"].value == "value"


def test_superblock_and_group_structures_follow_the_spec(tmp_path):
    fn = str(tmp_path / "s.h5")
    a = np.arange(24.0).reshape((4, 3, 2), order="F")
    hdf5.write_matrix(fn, "p0", a)
    b = open(fn, "rb").read()
    assert b[:8] == b"\x89HDF\r\n\x1a\n"
    assert b[8:16] == bytes([0, 0, 0, 0, 0, 8, 8, 0])        # versions, offset/length sizes
    assert struct.unpack_from("<HH", b, 16) == (4, 16)       # group leaf / internal K
    assert _u64(b, 24) == 0 and _u64(b, 32) == hdf5.UNDEF    # base address, free space
    assert _u64(b, 40) == len(b)                             # end of file address
    root_oh = _u64(b, 64)
    assert struct.unpack_from("<I", b, 72)[0] == 1           # cache type 1: B-tree + heap in scratch
    btree, heap = _u64(b, 80), _u64(b, 88)
    assert b[btree:btree + 4] == b"TREE" and b[btree + 4] == 0
    assert b[heap:heap + 4] == b"HEAP"
    # root object header (version 1) carries a symbol table message with the same addresses
    assert b[root_oh] == 1
    mtype, msize = struct.unpack_from("<HH", b, root_oh + 16)
    assert mtype == 0x11 and (_u64(b, root_oh + 24), _u64(b, root_oh + 32)) == (btree, heap)
    # B-tree leaf -> symbol table node -> dataset object header
    snod = _u64(b, btree + 32)
    assert b[snod:snod + 4] == b"SNOD" and struct.unpack_from("<H", b, snod + 6)[0] == 1
    name_off, ds_oh = _u64(b, snod + 8), _u64(b, snod + 16)
    data_seg = _u64(b, heap + 24)
    assert b[data_seg + name_off:data_seg + name_off + 3] == b"p0\0"
    # dataset: dataspace dims are the Julia shape reversed; chunk B-tree holds one zlib chunk
    msgs = {}
    p, n = ds_oh + 16, struct.unpack_from("<H", b, ds_oh + 2)[0]
    for _ in range(n):
        t, s = struct.unpack_from("<HH", b, p)
        msgs[t] = b[p + 8:p + 8 + s]
        p += 8 + s
    ds = msgs[0x0001]
    assert ds[0] == 1 and ds[1] == 3 and [_u64(ds, 8 + 8 * i) for i in range(3)] == [2, 3, 4]
    lay = msgs[0x0008]
    assert lay[:3] == bytes([3, 2, 4])                       # version 3, chunked, rank + 1
    assert struct.unpack_from("<4I", lay, 11) == (2, 3, 4, 8)
    ct = _u64(lay, 3)
    assert b[ct:ct + 4] == b"TREE" and b[ct + 4] == 1 and struct.unpack_from("<H", b, ct + 6)[0] == 1
    csize = struct.unpack_from("<I", b, ct + 24)[0]
    caddr = _u64(b, ct + 24 + 40)
    assert zlib.decompress(b[caddr:caddr + csize]) == a.ravel(order="F").tobytes()
    pipe = msgs[0x000B]
    assert pipe[0] == 1 and pipe[1] == 1 and struct.unpack_from("<H", pipe, 8)[0] == 1   # deflate
    assert struct.unpack_from("<I", pipe, 24)[0] == 3                                     # level 3


@pytest.mark.skipif(MAT73 is None, reason="scipy's HDF5-based MATLAB test file is not installed")
def test_reader_on_a_file_written_by_libhdf5():
    f = hdf5.read_file(MAT73)
    v = f.datasets["testdouble"].value
    assert v.shape == (1, 9)                                 # MATLAB 1x9, column-major like Julia
    np.testing.assert_allclose(v.ravel(), np.arange(9) * np.pi / 4, rtol=0, atol=1e-15)
    assert f.datasets["testdouble"].attrs["MATLAB_class"] == "double"


@pytest.mark.skipif(MAT73 is None, reason="scipy's HDF5-based MATLAB test file is not installed")
def test_datatype_encodings_match_libhdf5_bytes():
    raw = open(MAT73, "rb").read()
    want_f64 = bytes.fromhex("11203f00080000000000400034 0b003 4ff030000".replace(" ", ""))
    assert want_f64 in raw                                   # what libhdf5 wrote for a double
    assert hdf5._dtype_message(np.float64) == want_f64
    # a 6-byte ASCII string datatype in that file; ours differs only in the UTF-8 character set bit
    assert bytes.fromhex("1300000006000000") in raw
    assert hdf5._dtype_message("double") == bytes.fromhex("1310000006000000")


def test_write_grid_flags_attributes_and_read_output(tmp_path):
    fn = str(tmp_path / "g.h5")
    kg = K.KWaveGrid(32, 1e-4, 24, 2e-4)
    kg.dt, kg.Nt = 2e-8, 100
    hdf5.write_grid(fn, kg, pml_size=(10, 12), pml_alpha=(2.0, 1.5))
    hdf5.write_attributes(fn)
    out = hdf5.read_output(fn)
    assert list(out["Nx"]) == [32] and list(out["Ny"]) == [24]
    assert list(out["dx"]) == [1e-4] and list(out["dy"]) == [2e-4]
    assert list(out["Nt"]) == [100] and list(out["dt"]) == [2e-8]
    assert list(out["pml_x_size"]) == [10] and list(out["pml_y_size"]) == [12]
    assert list(out["pml_x_alpha"]) == [2.0] and list(out["pml_y_alpha"]) == [1.5]
    attrs = hdf5.read_file(fn).attrs
    assert attrs["created_by"] == "KWave.jl" and attrs["file_type"] == "input"
    assert attrs["major_version"] == "1" and attrs["minor_version"] == "2"


def _p0_problem(nd):
    if nd == 2:
        kg = K.KWaveGrid(32, 1e-4, 24, 1e-4)
        c0 = np.full((32, 24), 1500.0)
        c0[20:, :] = 1600.0
        p0 = np.zeros((32, 24))
        p0[14:18, 10:14] = 1.0
        med = K.KWaveMedium(sound_speed=c0, density=np.full((32, 24), 1000.0))
    else:
        kg = K.KWaveGrid(40, 1e-4)
        p0 = np.zeros(40)
        p0[18:22] = 1.0
        med = K.KWaveMedium(sound_speed=1500.0, density=1000.0)
    K.make_time(kg, 1600.0)
    return kg, med, K.KWaveSource(p0=p0), K.KWaveSensor(mask=np.ones(kg.grid_size(), dtype=bool), record=["p"])


def test_save_to_disk_writes_the_input_file(tmp_path):
    fn = str(tmp_path / "in.h5")
    kg, med, src, sen = _p0_problem(2)
    out = K.kspace_first_order(kg, med, src, sen, save_to_disk=fn, pml_size=(8, 6), pml_alpha=2.0)
    assert len(list(out.keys())) == 0
    f = hdf5.read_file(fn)
    assert set(f.datasets) == {"Nx", "Ny", "dx", "dy", "Nt", "dt", "pml_x_size", "pml_y_size", "pml_x_alpha",
                               "pml_y_alpha", "c0", "rho0", "p0"}
    np.testing.assert_array_equal(f.datasets["c0"].value, med.sound_speed)
    np.testing.assert_array_equal(f.datasets["p0"].value, src.p0)
    assert list(f.datasets["pml_y_size"].value) == [6]
    # 1-D: scalar c0 goes out as c_ref and rho0 is not written (first_order.jl:129-140)
    fn1 = str(tmp_path / "in1.h5")
    kg, med, src, sen = _p0_problem(1)
    K.kspace_first_order(kg, med, src, sen, save_to_disk=fn1)
    f1 = hdf5.read_file(fn1)
    assert "c_ref" in f1.datasets and "c0" not in f1.datasets and "rho0" not in f1.datasets
    assert list(f1.datasets["c_ref"].value) == [1500.0]
    # validation runs before the file is written
    kg.Nt = 0
    with pytest.raises(ValueError, match="Time array not set"):
        K.kspace_first_order(kg, med, src, sen, save_to_disk=str(tmp_path / "never.h5"))
    assert not os.path.exists(str(tmp_path / "never.h5"))


def test_cli_validate_info_and_errors(tmp_path, capsys):
    fn = str(tmp_path / "in.h5")
    kg, med, src, sen = _p0_problem(2)
    K.kspace_first_order(kg, med, src, sen, save_to_disk=fn)
    assert cli.main(["validate", "--input", fn]) == 0
    o = capsys.readouterr().out
    assert "Valid k-Wave input file" in o and "Nx = 32" in o and "Ny = 24" in o and f"Nt = {kg.Nt}" in o
    assert cli.main(["info", "-i", fn]) == 0
    o = capsys.readouterr().out
    assert "created_by = KWave.jl" in o and "c0: Float64 (32, 24)" in o and "Nx: Int64 (1,)" in o
    assert cli.main([]) == 1
    assert cli.main(["help"]) == 0
    assert cli.main(["bogus"]) == 1
    assert cli.main(["run", "--output", "x.h5"]) == 1
    assert cli.main(["run", "--input", fn]) == 1
    assert cli.main(["run", "--input", str(tmp_path / "nope.h5"), "-o", "x.h5"]) == 1
    assert cli.main(["validate"]) == 1
    bad = str(tmp_path / "bad.h5")
    hdf5.write_matrix(bad, "Nx", 8)
    assert cli.main(["validate", "--input", bad]) == 1
    assert "missing fields: dx, Nt, dt" in capsys.readouterr().err
