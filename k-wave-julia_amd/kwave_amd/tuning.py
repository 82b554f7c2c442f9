"""Launch-shape autotuner front end (kw_autotune / kw_tune_*).

The reference tunes its transforms with FFTW.MEASURE when it creates its
plans (fft_utils.jl:44-62).  Here each pass is a hand-written line-FFT
kernel whose one free launch parameter is the number of lines a workgroup
carries; ``autotune`` times the candidates on the device for a given
simulation and keeps the fastest per (precision, pass, line length) in the
library's process-wide table, which every later simulation or batch reads.
The table can be saved to and loaded from a small text cache file.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Sequence, Tuple

import numpy as np

from . import _lib
from .solver import build_desc

PASS_NAMES = ("dir_x", "dir_y", "dir_z", "spec_y", "spec_z", "c2r_x")


def autotune(kgrid, medium, source, sensor, *, steps: int = 10, data_cast=np.float32, device: int = 0,
             cache: str | None = None, **kwargs) -> Dict[Tuple[int, str, int], int]:
    """Time the launch shapes of this simulation's passes; returns the table.

    Keyword arguments are those of ``kspace_first_order``.  With ``cache`` the
    table is loaded from that file first (if present) and written back after.
    """
    lib = _lib.load()
    if cache and os.path.exists(cache):
        load_table(cache)
    desc, keep, _ = build_desc(kgrid, medium, source, sensor, data_cast=data_cast, **kwargs)
    _lib.check(lib.kw_autotune(_lib.context(device), C.byref(desc), 1, int(steps)))
    del keep
    if cache:
        save_table(cache)
    return table()


def autotune_batch(sims: Sequence[Tuple], *, steps: int = 10, data_cast=np.float32, device: int = 0,
                   cache: str | None = None, **kwargs) -> Dict[Tuple[int, str, int], int]:
    """Time the launch shapes of a batch of (kgrid, medium, source, sensor) run together.

    The trials build and run the whole batch, so the table holds the shapes that
    are fastest when the simulations share each launch (kspace_first_order_batch).
    """
    lib = _lib.load()
    if cache and os.path.exists(cache):
        load_table(cache)
    descs, keep = [], []
    for kgrid, medium, source, sensor in sims:
        d, k, _ = build_desc(kgrid, medium, source, sensor, data_cast=data_cast, **kwargs)
        descs.append(d)
        keep.append(k)
    arr = (_lib.SimDesc * len(descs))(*descs)
    _lib.check(lib.kw_autotune(_lib.context(device), arr, len(descs), int(steps)))
    del keep
    if cache:
        save_table(cache)
    return table()


def table_text() -> str:
    lib = _lib.load()
    n = C.c_int64()
    _lib.check(lib.kw_tune_export(None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    _lib.check(lib.kw_tune_export(buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def table() -> Dict[Tuple[int, str, int], int]:
    """{(precision bytes, pass name, line length): lines per workgroup}."""
    out = {}
    for line in table_text().splitlines():
        p, c, n, l = (int(x) for x in line.split())
        out[(p, PASS_NAMES[c], n)] = l
    return out


def load_table(path: str) -> None:
    with open(path) as f:
        _lib.check(_lib.load().kw_tune_import(f.read().encode()))


def save_table(path: str) -> None:
    with open(path, "w") as f:
        f.write(table_text())


def clear_table() -> None:
    _lib.check(_lib.load().kw_tune_clear())
