"""Tile-size autotuner for the fused stage kernels and the batched solver.

Candidates are timed on the device with events on the launch stream; the
winner is cached per (dtype, flux, velocity mode, reconstruction, size class) for the stage
kernels and per (dtype, flux) for the batched solver, and used by
:func:`~pyshocks_amd.timestepping.fused_plan` and
:class:`~pyshocks_amd.batched.BatchedSolver` when no tiles are given. A
candidate the device cannot launch is skipped, and so is a batched kernel
that spills registers to private memory.
"""

from __future__ import annotations

import math
from typing import Any, Sequence

import numpy as np
import torch

STAGE_CANDIDATES: tuple[tuple[int, int], ...] = ((256, 4), (256, 2), (512, 2), (128, 4), (256, 1))
BATCH_CANDIDATES: tuple[tuple[int, int], ...] = ((256, 4), (256, 2), (512, 2), (512, 4), (1024, 2))

_STAGE_CACHE: dict[tuple[Any, ...], tuple[int, int]] = {}
_BATCH_CACHE: dict[tuple[Any, ...], tuple[int, int]] = {}


def _stage_key(desc: Any) -> tuple[Any, ...]:
    vel = 0 if desc.flux != 2 else (1 if desc.velocity_is_const else 2)
    return (int(desc.dtype), int(desc.flux), vel, bool(desc.mlp), int(desc.rec),
            max(0, math.ceil(math.log2(max(desc.n, 1)))))


def lookup_stage(desc: Any) -> tuple[int, int] | None:
    return _STAGE_CACHE.get(_stage_key(desc))


def tune_stage(plan: Any, u0: torch.Tensor, *, steps: int = 4,
               candidates: Sequence[tuple[int, int]] = STAGE_CANDIDATES,
               ) -> tuple[tuple[int, int], dict[tuple[int, int], float]]:
    """Time ``steps`` fused steps from ``u0`` for each candidate and cache
    the fastest; returns ``(best, {candidate: ms per step})``."""
    dev = u0.device
    u = u0.clone()
    k1 = torch.empty_like(u)
    k2 = torch.empty_like(u)
    old = (plan.desc.block, plan.desc.cells_per_thread)
    times: dict[tuple[int, int], float] = {}
    tmax = float(np.finfo(plan.np_dtype).max)
    for cand in candidates:
        plan.desc.block, plan.desc.cells_per_thread = cand
        try:
            u.copy_(u0)
            ctrl = plan.new_control(u, 0.0, tmax, None)
            plan.steps(u, u, k1, k2, ctrl, 1)  # warm up
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            plan.steps(u, u, k1, k2, ctrl, steps)
            e1.record()
            torch.cuda.synchronize(dev)
            times[cand] = e0.elapsed_time(e1) / steps
        except (RuntimeError, ValueError):
            continue
    plan.desc.block, plan.desc.cells_per_thread = old
    if not times:
        raise RuntimeError("no tile candidate could be launched")
    best = min(times, key=lambda c: times[c])
    _STAGE_CACHE[_stage_key(plan.desc)] = best
    return best, times


def lookup_batch(dtype: int, flux: int, learned: bool = False) -> tuple[int, int] | None:
    return _BATCH_CACHE.get((int(dtype), int(flux), bool(learned)))


def tune_batch(ncells: Sequence[int], states: Sequence[np.ndarray], *,
               dtype: torch.dtype = torch.float32, scheme: str = "rusanov",
               tfinal: float = 0.01, eps: float = 1.0e-6, theta: float = 1.0,
               candidates: Sequence[tuple[int, int]] = BATCH_CANDIDATES,
               mlp: Any = None, rec: Any = None, team_cells: int | None = None,
               ) -> tuple[tuple[int, int], dict[tuple[int, int], float]]:
    """Time the batched solve of the instances (to a short final time) for
    each candidate and cache the fastest; with a reconstruction other than
    WENO-JS5 (``rec``, states with 2 ghost cells) the winner is returned but
    not cached. A candidate whose kernel spills registers to private memory
    is skipped without being timed. Each candidate runs with the teams the
    solver gives these instances under ``team_cells`` (``None``: the
    solver's default), which depend on the candidate's resident workgroups:
    passed the whole batch, the tuner picks tiles and teams together, and a
    short final time keeps the batch's imbalance, since every instance's
    step count scales with it."""
    from pyshocks_amd import _lib
    from pyshocks_amd.batched import BatchedSolver

    times: dict[tuple[int, int], float] = {}
    for cand in candidates:
        try:
            solver = BatchedSolver(ncells, dtype=dtype, scheme=scheme, eps=eps,
                                   theta=theta, tiles=cand, mlp=mlp, rec=rec,
                                   nghosts=3 if rec is None else 2, team_cells=team_cells)
            if solver.kernel_info()["private_bytes"] > 0:
                continue
            solver.set_states(states)
            u0 = solver.u.clone()
            solver.solve(tfinal)  # warm up
            solver.u.copy_(u0)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            solver.solve(tfinal)
            e1.record()
            torch.cuda.synchronize()
            times[cand] = e0.elapsed_time(e1)
        except (RuntimeError, ValueError):
            continue
    if not times:
        raise RuntimeError("no tile candidate could be launched")
    best = min(times, key=lambda c: times[c])
    flux = _lib.BATCH_FLUX[scheme]
    if rec is not None:
        return best, times
    _BATCH_CACHE[(_lib.dtype_code(dtype), flux, mlp is not None)] = best
    return best, times
