"""Gradient of a sensor misfit with respect to the sound-speed map.

The reference differentiates ``kspace_first_order`` (first_order.jl:505-728)
only through a generic AD tool.  Here the gradient comes from the discrete
adjoint of the GPU time step (``kw_sim_grad_c0``): the forward run keeps
state checkpoints every ``checkpoint_every`` steps, the backward sweep
recomputes each segment from its checkpoint and runs the transposed step.

    loss = 1/2 sum_{firings} sum_{j,t} (p_j(t) - p_obs_j(t))^2

``gradient_c0`` returns (loss, dL/dc0); ``misfit_rrule`` returns the loss and
a pullback, the shape a ChainRulesCore ``rrule`` has on the Julia side.
"""
from __future__ import annotations

from typing import Callable, Optional, Sequence, Tuple, Union

import numpy as np

from .model import KWaveMedium
from .solver import Simulation, build_desc


def _firings(source, sensor, p_obs):
    if isinstance(source, (list, tuple)):
        sensors = sensor if isinstance(sensor, (list, tuple)) else [sensor] * len(source)
        if not (len(source) == len(sensors) == len(p_obs)):
            raise ValueError("sources, sensors and observations must have the same length")
        return list(zip(source, sensors, p_obs))
    return [(source, sensor, p_obs)]


def gradient_c0(kgrid, medium: KWaveMedium, source, sensor, p_obs, *, checkpoint_every: int = 50,
                data_cast=np.float32, device: int = 0, stack: Optional[int] = None,
                **kwargs) -> Tuple[float, np.ndarray]:
    """Misfit and its gradient w.r.t. ``medium.sound_speed`` (a map).

    ``source``/``sensor``/``p_obs`` may be lists (one entry per firing; the
    losses and gradients are summed).  Each ``p_obs`` is n_sensor x Nt.
    Keyword arguments are those of ``kspace_first_order`` (pml_size,
    pml_alpha, smooth_p0, ...).  kgrid.dt and kgrid.Nt are used as set.

    On 2-D grids the firings run stacked (kw_sim_create_firings): every pass
    of the forward, recompute and adjoint sweeps covers ``stack`` firings at
    once (default: all of them); ``stack=1`` runs them one after another.
    """
    c0 = np.asarray(medium.sound_speed, dtype=np.float64)
    if c0.ndim == 0:
        raise ValueError("gradient_c0 needs medium.sound_speed as a map")
    firings = _firings(source, sensor, p_obs)
    prepared = []
    for src, sen, obs in firings:
        if "p" not in sen.record:
            raise ValueError("the misfit needs 'p' in sensor.record")
        desc, keep, info = build_desc(kgrid, medium, src, sen, data_cast=data_cast, **kwargs)
        obs = np.asarray(obs, dtype=np.float64)
        if obs.shape != (info["n_sensor"], info["Nt"]):
            raise ValueError(f"p_obs must be {(info['n_sensor'], info['Nt'])}, got {obs.shape}")
        prepared.append((desc, keep, info, obs, src))
    stacked = kgrid.ndim == 2 and len(prepared) > 1 and (stack is None or stack > 1) and all(
        x[4].p0 is None for x in prepared)
    total, grad = 0.0, np.zeros(c0.shape)
    if not stacked:
        for desc, keep, info, obs, _ in prepared:
            with Simulation(desc, keep, device=device) as s:
                loss, g = s.grad_c0(obs, checkpoint_every)
            total += loss
            grad += g
        return total, grad
    size = len(prepared) if stack is None else int(stack)
    for i in range(0, len(prepared), size):
        group = prepared[i:i + size]
        obs = np.concatenate([x[3] for x in group], axis=0)
        with Simulation.firings([x[0] for x in group], [x[1] for x in group], device=device) as s:
            loss, g = s.grad_c0(obs, checkpoint_every)
        total += loss
        grad += g
    return total, grad


def misfit_rrule(kgrid, medium: KWaveMedium, source, sensor, p_obs, **kwargs
                 ) -> Tuple[float, Callable[[float], np.ndarray]]:
    """(loss, pullback) with pullback(dl) = dl * dL/dc0."""
    loss, g = gradient_c0(kgrid, medium, source, sensor, p_obs, **kwargs)
    return loss, (lambda dl: float(dl) * g)
