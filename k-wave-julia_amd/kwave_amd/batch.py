"""Batched simulations: many independent kspace_first_order runs per launch.

The reference runs one simulation per call; a sweep over transmit elements
(one firing per simulation) is a loop of calls.  ``kspace_first_order_batch``
takes that list and runs it through ``kw_batch_*``: each pass of the time
step is one launch per line length over all simulations, the step index
lives on the device, and one step is captured in a hipGraph and replayed.
With ``persistent`` every batched pass runs as a persistent grid (as many
workgroups as fit the device at once, each looping over the pass's tiles).
Results are the same ``SimulationOutput`` objects the single call returns.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence, Tuple

import numpy as np

from . import _lib
from .model import SimulationOutput
from .solver import _DP, build_desc


class Batch:
    """Handle over kw_batch_create .. kw_batch_destroy."""

    GRAPH, PERSISTENT = 1, 2   # kw_batch_create flags (KW_BATCH_GRAPH, KW_BATCH_PERSISTENT)

    def __init__(self, descs: Sequence["_lib.SimDesc"], keep: list, use_graph: bool = True, device: int = 0,
                 persistent: bool = False, grid8: int = 0):
        self._lib = _lib.load()
        ctx = _lib.context(device)
        arr = (_lib.SimDesc * len(descs))(*descs)
        h = C.c_void_p()
        # grid8 > 0 (persistent only): at most 8 * grid8 workgroups per batched launch
        flags = (self.GRAPH if use_graph else 0) | (self.PERSISTENT if persistent else 0) | (int(grid8) << 8)
        _lib.check(self._lib.kw_batch_create(ctx, arr, len(descs), flags, C.byref(h)))
        self._h = h
        self.descs = list(descs)
        del keep

    def run(self, n_steps: int) -> None:
        _lib.check(self._lib.kw_batch_run(self._h, int(n_steps)))

    def synchronize(self) -> None:
        _lib.check(self._lib.kw_batch_synchronize(self._h))

    def sensor(self, i: int, name: str) -> np.ndarray:
        bit = _lib.REC_BITS[name]
        r, c = C.c_int64(), C.c_int64()
        _lib.check(self._lib.kw_batch_sensor_shape(self._h, int(i), bit, C.byref(r), C.byref(c)))
        out = np.empty(r.value * c.value, dtype=np.float64)
        _lib.check(self._lib.kw_batch_get_sensor(self._h, int(i), bit, out.ctypes.data_as(_DP)))
        return out if c.value == 1 else out.reshape((r.value, c.value), order="F")

    def field(self, i: int, name: str) -> np.ndarray:
        d = self.descs[i]
        shape = tuple(d.N[a] for a in range(d.ndim))
        out = np.empty(int(np.prod(shape)), dtype=np.float64)
        _lib.check(self._lib.kw_batch_get_field(self._h, int(i), _lib.FIELD[name], out.ctypes.data_as(_DP)))
        return out.reshape(shape, order="F")

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.check(self._lib.kw_batch_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kspace_first_order_batch(sims: Sequence[Tuple], *, data_cast=np.float32, use_graph: bool = True,
                             device: int = 0, autotune: bool = False, tune_cache: str | None = None,
                             persistent: bool = False, **kwargs) -> List[SimulationOutput]:
    """Run a list of (kgrid, medium, source, sensor) in one batch.

    Keyword arguments are those of ``kspace_first_order`` and apply to every
    simulation; all simulations must share dimensionality and Nt.  With
    ``autotune`` the launch shapes are first timed on the whole batch
    (``tune_cache``: text file the table is kept in); ``persistent`` selects
    the persistent-grid launches.
    """
    if autotune:
        from .tuning import autotune_batch
        autotune_batch(sims, data_cast=data_cast, device=device, cache=tune_cache, **kwargs)
    descs, keep, infos = [], [], []
    for kgrid, medium, source, sensor in sims:
        d, k, info = build_desc(kgrid, medium, source, sensor, data_cast=data_cast, **kwargs)
        if info["tr"]:
            raise NotImplementedError("time reversal is not batched")
        descs.append(d)
        keep.append(k)
        infos.append(info)
    outs = []
    with Batch(descs, keep, use_graph=use_graph, device=device, persistent=persistent) as b:
        b.run(infos[0]["Nt"])
        b.synchronize()
        for i, (kgrid, medium, source, sensor) in enumerate(sims):
            data = {}
            if sensor.mask is not None:
                for f in sensor.record:
                    if f in ("ux", "uy", "uz") and "xyz".index(f[1]) >= kgrid.ndim:
                        data[f] = np.zeros((infos[i]["n_sensor"], infos[i]["Nt"]))
                        continue
                    data[f] = b.sensor(i, f)
            outs.append(SimulationOutput(data))
    return outs
