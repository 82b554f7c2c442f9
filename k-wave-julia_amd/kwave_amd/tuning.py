"""Launch-shape autotuner front end (kw_autotune / kw_tune_*).

The reference tunes its transforms with FFTW.MEASURE when it creates its
plans (fft_utils.jl:44-62).  Here each pass is a hand-written line-FFT
kernel with two free choices: the number of lines a workgroup carries and,
for line lengths 256 and 512, the plan family (the radix schedule and
threads per line: 0 = the strided passes' plan, 1 = the contiguous x pass's,
2 = the complex-to-real pass's).  ``autotune`` times the candidates on the
device for a given simulation and keeps the fastest per (precision, pass,
line length) in the library's process-wide table, which every later single
simulation reads (batches take the lines per workgroup and the default plan
families).  The table can be saved to and loaded from a small text cache
file: one entry per line, ``precision class length lines [plan]``.
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


def _entries():
    for line in table_text().splitlines():
        v = [int(x) for x in line.split()]
        yield (v[0], PASS_NAMES[v[1]], v[2]), v[3], (v[4] if len(v) == 5 else None)


def table() -> Dict[Tuple[int, str, int], int]:
    """{(precision bytes, pass name, line length): lines per workgroup}."""
    return {k: l for k, l, _ in _entries()}


def plans() -> Dict[Tuple[int, str, int], int]:
    """{(precision bytes, pass name, line length): plan family} for the entries
    that name one (the others use the pass's default family)."""
    return {k: v for k, _, v in _entries() if v is not None}


def load_table(path: str) -> None:
    with open(path) as f:
        _lib.check(_lib.load().kw_tune_import(f.read().encode()))


def save_table(path: str) -> None:
    with open(path, "w") as f:
        f.write(table_text())


def clear_table() -> None:
    _lib.check(_lib.load().kw_tune_clear())
