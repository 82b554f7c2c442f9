"""Build libkwave_hip.so for gfx950 from the HIP sources in this folder.

Each translation unit is compiled in its own hipcc process (in parallel),
then linked into ``../kwave_amd/libkwave_hip.so``.  Usage:

    python k-wave-julia_amd/csrc/build.py [-j JOBS]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "kwave_amd")
OUT = os.path.join(PKG, "libkwave_hip.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = [
    "kw_launch_dir_f32.hip",
    "kw_launch_dir_f64.hip",
    "kw_launch_spec_f32.hip",
    "kw_launch_spec_f64.hip",
    "kw_launch_c2r_f32.hip",
    "kw_launch_c2r_f64.hip",
    "kw_launch_batch_f32.hip",
    "kw_launch_batch_f64.hip",
    "kw_launch_adj.hip",
    "kw_launch_misc.hip",
    "kw_launch_el_f32.hip",
    "kw_launch_el_f64.hip",
    "kw_launch_bs_f32.hip",
    "kw_launch_bs_f64.hip",
    "kw_api.hip",
    "kw_elastic.hip",
    "kw_ops.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for every device compile.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-I" + HERE,
         "-I" + os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")] + DEVICE_FLAGS


def _deps(path: str, seen=None) -> set:
    """The file and every quoted #include it reaches (recursively)."""
    seen = set() if seen is None else seen
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s.startswith("#include \""):
                inc = s.split("\"")[1]
                _deps(os.path.normpath(os.path.join(os.path.dirname(path), inc)), seen)
    return seen


def _stale(src: str, obj: str) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in _deps(os.path.join(HERE, src)))


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(HERE, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
    return obj


def build(jobs: int = 0, verbose: bool = True, incremental: bool = False) -> str:
    """Compile every translation unit (incremental: only those whose source or
    included headers changed since their object was built) and link."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} in {time.time() - t0:.1f}s", flush=True)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=0)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only translation units whose sources or headers changed")
    a = ap.parse_args()
    try:
        build(a.jobs, incremental=a.incremental)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
