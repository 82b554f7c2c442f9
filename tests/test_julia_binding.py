"""The Julia drop-in (julia/KWaveHIP) against the C ABI, without Julia.

The image has no Julia, so the package is checked as text: its KwSimDesc
mirrors kw_sim_desc field for field (names, order, C types; Julia lays out an
isbits struct with C alignment, so equal field types give equal offsets), and
every library symbol it calls is declared in include/kwave_hip.h with the
same number of arguments.
"""
import ctypes as C
import os
import re

from kwave_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "julia", "KWaveHIP")
HEADER = os.path.join(ROOT, "include", "kwave_hip.h")

_JL_TO_C = {"Int32": C.c_int32, "Int64": C.c_int64, "UInt32": C.c_uint32, "Float64": C.c_double,
            "Ptr{Float64}": C.POINTER(C.c_double), "Ptr{Int64}": C.POINTER(C.c_int64)}


def _jl_sources():
    out = {}
    for d in ("src", "ext"):
        for f in sorted(os.listdir(os.path.join(PKG, d))):
            if f.endswith(".jl"):
                out[f] = open(os.path.join(PKG, d, f)).read()
    return out


def _jl_struct_fields(src, name):
    body = re.search(r"struct\s+" + name + r"\b(.*?)\nend", src, flags=re.S).group(1)
    fields = []
    for line in body.splitlines():
        m = re.match(r"\s*(\w+)::([^=#]+?)\s*(=.*)?$", line)
        if m:
            fields.append((m.group(1), m.group(2).strip()))
    return fields


def _ctype(jl):
    m = re.fullmatch(r"NTuple\{(\d+),\s*(.+)\}", jl)
    if m:
        return _JL_TO_C[m.group(2).strip()] * int(m.group(1))
    return _JL_TO_C[jl]


def test_kwsimdesc_mirrors_the_c_descriptor():
    src = _jl_sources()["KWaveHIP.jl"]
    jl = _jl_struct_fields(src, "KwSimDesc")
    ct = _lib.SimDesc._fields_
    assert [n for n, _ in jl] == [n for n, _ in ct]
    for (n, t), (_, c) in zip(jl, ct):
        want = _ctype(t)
        if hasattr(c, "_length_"):
            assert hasattr(want, "_length_") and want._length_ == c._length_, n
            assert C.sizeof(want._type_) == C.sizeof(c._type_), n
            assert want._type_.__name__ == c._type_.__name__ or (
                issubclass(want._type_, C._Pointer) and issubclass(c._type_, C._Pointer)), n
        else:
            assert C.sizeof(want) == C.sizeof(c), n
            assert want.__name__ == c.__name__ or (issubclass(want, C._Pointer) and issubclass(c, C._Pointer)), n


def _header_arity():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(kw_\w+)\s*\(([^;{]*?)\)\s*;", txt, flags=re.S):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else len(args.split(","))
    return out


def test_every_ccall_is_declared_with_its_arity():
    arity = _header_arity()
    seen = set()
    for name, src in _jl_sources().items():
        for m in re.finditer(r"ccall\(\s*sym\(:(kw_\w+)\)\s*,\s*\w+\s*,\s*\(([^()]*(?:\([^()]*\)[^()]*)*)\)", src):
            fn, types = m.group(1), m.group(2).strip().rstrip(",")
            n = 0 if not types else len([t for t in re.split(r",(?![^{]*\})", types) if t.strip()])
            assert fn in arity, (name, fn)
            assert arity[fn] == n, (name, fn, arity[fn], n)
            seen.add(fn)
    for fn in ("kw_ctx_create", "kw_sim_create", "kw_sim_run", "kw_sim_get_sensor", "kw_sim_sensor_shape",
               "kw_sim_get_field", "kw_sim_destroy", "kw_sim_grad_c0", "kw_last_error"):
        assert fn in seen, fn


def test_package_layout():
    proj = open(os.path.join(PKG, "Project.toml")).read()
    assert 'name = "KWaveHIP"' in proj
    # the extension loads on the reference package's own uuid (KWave.jl/Project.toml)
    assert 'KWave = "6515db03-cc09-4e48-9b6f-9f8a641b1330"' in proj
    src = _jl_sources()
    assert "KWaveHIPKWaveExt.jl" in src and "KWaveHIPChainRulesCoreExt.jl" in src
    # the reference's keywords, all accepted
    sig = re.search(r"function _gpu_kspace_first_order\((.*?)\)\n", src["KWaveHIP.jl"], flags=re.S).group(1)
    for kw in ("pml_inside", "pml_size", "pml_alpha", "smooth_p0", "smooth_c0", "smooth_rho0", "data_cast",
               "save_to_disk", "plot_sim", "plot_layout", "plot_scale", "record_movie", "progress_callback",
               "show_progress"):
        assert re.search(r"\b" + kw + r"\b", sig), kw
