"""Inviscid Burgers' equation: flux, time step, Godunov and Rusanov schemes.

Mirrors ``burgers/schemes.py`` of pyshocks (``BurgersScheme`` ``:30-34``,
``flux`` ``:37-39``, ``predict_timestep`` ``:42-49``, ``FiniteVolumeScheme``
``:53-58``, ``Godunov`` ``:75-89``, ``Rusanov`` ``:99-127``) and the registry
of ``burgers/__init__.py:62-98``.
"""

from __future__ import annotations

import ctypes

from dataclasses import dataclass, fields
from typing import TYPE_CHECKING, Any

import torch

from pyshocks_amd import _lib
from pyshocks_amd.reductions import max_abs_interior
from pyshocks_amd.schemes import (
    Boundary,
    ConservationLawScheme,
    SchemeBase,
    apply_operator,
    flux,
    numerical_flux,
    predict_timestep,
)

if TYPE_CHECKING:
    from pyshocks_amd.grid import Grid


@dataclass(frozen=True)
class BurgersScheme(SchemeBase):
    """Base class for numerical schemes for Burgers' equation."""


@flux.register(BurgersScheme)
def _flux_burgers(scheme: BurgersScheme, t: Any, x: torch.Tensor,
                  u: torch.Tensor) -> torch.Tensor:
    return u**2 / 2


@predict_timestep.register(BurgersScheme)
def _predict_timestep_burgers(scheme: BurgersScheme, grid: Grid, bc: Boundary,
                              t: Any, u: torch.Tensor) -> torch.Tensor:
    smax = max_abs_interior(u, grid.nghosts)
    return 0.5 * grid.dx_min / smax


@dataclass(frozen=True)
class FiniteVolumeScheme(BurgersScheme, ConservationLawScheme):
    """Finite volume schemes for Burgers' equation."""


@dataclass(frozen=True)
class Godunov(FiniteVolumeScheme):
    """Upwind flux :func:`~pyshocks_amd.scalar.scalar_flux_upwind`."""


@numerical_flux.register(Godunov)
def _numerical_flux_burgers_godunov(scheme: Godunov, grid: Grid, bc: Boundary,
                                    t: Any, u: torch.Tensor) -> torch.Tensor:
    from pyshocks_amd.scalar import scalar_flux_upwind

    return scalar_flux_upwind(scheme, grid, bc.boundary_type, t, u, u)


@dataclass(frozen=True)
class Rusanov(FiniteVolumeScheme):
    """Rusanov (local Lax-Friedrichs) flux
    :func:`~pyshocks_amd.scalar.scalar_flux_rusanov`."""

    alpha: float = 1.0


@numerical_flux.register(Rusanov)
def _numerical_flux_burgers_rusanov(scheme: Rusanov, grid: Grid, bc: Boundary,
                                    t: Any, u: torch.Tensor) -> torch.Tensor:
    from pyshocks_amd.scalar import scalar_flux_rusanov

    return scalar_flux_rusanov(scheme, grid, bc.boundary_type, t, u, u,
                               alpha=scheme.alpha)


@predict_timestep.register(Rusanov)
def _predict_timestep_burgers_rusanov(scheme: Rusanov, grid: Grid, bc: Boundary,
                                      t: Any, u: torch.Tensor) -> torch.Tensor:
    smax = max_abs_interior(u, grid.nghosts)
    return 0.5 * grid.dx_min ** (2 - scheme.alpha) / smax


def flux_kind(scheme: FiniteVolumeScheme) -> int:
    if isinstance(scheme, Rusanov):
        return _lib.PSA_FLUX_RUSANOV
    if isinstance(scheme, Godunov):
        return _lib.PSA_FLUX_BURGERS_UPWIND
    raise NotImplementedError(type(scheme).__name__)


@apply_operator.register(FiniteVolumeScheme)
def _apply_operator_burgers(scheme: FiniteVolumeScheme, grid: Grid, bc: Boundary,
                            t: Any, u: torch.Tensor) -> torch.Tensor:
    """``schemes.py:339-346`` fused: periodic ghost fill, WENO-JS5 faces and
    the conservative difference in one kernel."""
    from pyshocks_amd.scalar import PeriodicBoundary, _eps_of, rusanov_nu

    if not isinstance(bc, PeriodicBoundary):
        raise NotImplementedError(
            f"boundary {type(bc).__name__} is not on the GPU path")
    nu = rusanov_nu(grid, scheme.alpha) if isinstance(scheme, Rusanov) else None
    _lib.require_device(u, nu)
    out = torch.empty_like(u)
    from pyshocks_amd.reconstruction import LearnedWENOJS53

    if isinstance(scheme.rec, LearnedWENOJS53):
        # the operator stage of the descriptor carries the learned weights
        from pyshocks_amd.timestepping import BoundOperator, _scheme_plan

        sp = _scheme_plan(BoundOperator(scheme, grid, bc))
        assert sp is not None
        desc, _keep = sp
        _lib.call("psa_ssprk33_stage", ctypes.byref(desc), 0, None, _lib.ptr(u),
                  _lib.ptr(out), None, None, _lib.stream_of(u))
        return out
    _lib.call("psa_apply_operator", _lib.dtype_code(u.dtype), flux_kind(scheme),
              _lib.ptr(u), None, _lib.ptr(nu), _lib.ptr(out), u.numel(),
              grid.nghosts, grid.h, _eps_of(scheme), _lib.stream_of(u))
    return out


# {{{ registry

_SCHEMES: dict[str, type[BurgersScheme]] = {
    "godunov": Godunov,
    "rusanov": Rusanov,
}

_NOT_ON_GPU = ("default", "lf", "eo", "esweno32", "ssmuscl", "ssweno242",
               "splitrusanov")


def scheme_ids() -> tuple[str, ...]:
    return tuple(_SCHEMES.keys())


def make_scheme_from_name(name: str, **kwargs: Any) -> BurgersScheme:
    """``burgers/__init__.py:82-98``: unknown fields are ignored."""
    cls = _SCHEMES.get(name)
    if cls is None:
        if name in _NOT_ON_GPU:
            raise NotImplementedError(
                f"Scheme {name!r} is not part of the GPU hot path; available: "
                f"{', '.join(scheme_ids())}.")
        raise ValueError(f"Scheme {name!r} not found. Try one of "
                         f"{', '.join(scheme_ids())}.")

    return cls(**{f.name: kwargs[f.name] for f in fields(cls) if f.name in kwargs})


# }}}
