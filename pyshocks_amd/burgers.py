"""Inviscid Burgers' equation: flux, time step, Godunov, Rusanov,
Lax-Friedrichs and Engquist-Osher schemes.

Mirrors ``burgers/schemes.py`` of pyshocks (``BurgersScheme`` ``:30-34``,
``flux`` ``:37-39``, ``predict_timestep`` ``:42-49``, ``FiniteVolumeScheme``
``:53-58``, ``Godunov`` ``:75-89``, ``Rusanov`` ``:99-127``,
``LaxFriedrichs`` ``:136-153``, ``EngquistOsher`` ``:162-196``, ``ESWENO32``
``:202-256``) and the registry of ``burgers/__init__.py:62-98``.
"""

from __future__ import annotations

import ctypes

from dataclasses import dataclass, field, fields, replace
from typing import TYPE_CHECKING, Any

import torch

from pyshocks_amd import _lib
from pyshocks_amd.reductions import max_abs_interior
from pyshocks_amd.schemes import (
    Boundary,
    ConservationLawScheme,
    SchemeBase,
    apply_operator,
    bind,
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


@dataclass(frozen=True)
class LaxFriedrichs(Rusanov):
    """Lax-Friedrichs flux :func:`~pyshocks_amd.scalar.scalar_flux_lax_friedrichs`:
    Rusanov with the global speed ``max|u|`` (and Rusanov's time step)."""


@numerical_flux.register(LaxFriedrichs)
def _numerical_flux_burgers_lax_friedrichs(scheme: LaxFriedrichs, grid: Grid,
                                           bc: Boundary, t: Any,
                                           u: torch.Tensor) -> torch.Tensor:
    from pyshocks_amd.scalar import scalar_flux_lax_friedrichs

    return scalar_flux_lax_friedrichs(scheme, grid, bc.boundary_type, t, u, u,
                                      alpha=scheme.alpha)


@dataclass(frozen=True)
class EngquistOsher(FiniteVolumeScheme):
    """Engquist-Osher flux ``f(max(u_l, omega)) + f(min(u_r, omega)) - f(omega)``
    (:func:`~pyshocks_amd.scalar.scalar_flux_engquist_osher`) with ``omega = 0``,
    where the Burgers flux has its minimum."""

    omega: float = field(default=0, init=False, repr=False)


@numerical_flux.register(EngquistOsher)
def _numerical_flux_burgers_engquist_osher(scheme: EngquistOsher, grid: Grid,
                                           bc: Boundary, t: Any,
                                           u: torch.Tensor) -> torch.Tensor:
    from pyshocks_amd.scalar import scalar_flux_engquist_osher

    return scalar_flux_engquist_osher(scheme, grid, bc.boundary_type, t, u, u,
                                      omega=scheme.omega)


def flux_kind(scheme: FiniteVolumeScheme) -> int:
    # LaxFriedrichs before its base class Rusanov
    if isinstance(scheme, LaxFriedrichs):
        return _lib.PSA_FLUX_LAX_FRIEDRICHS
    if isinstance(scheme, Rusanov):
        return _lib.PSA_FLUX_RUSANOV
    if isinstance(scheme, Godunov):
        return _lib.PSA_FLUX_BURGERS_UPWIND
    if isinstance(scheme, EngquistOsher):
        return _lib.PSA_FLUX_ENGQUIST_OSHER
    raise NotImplementedError(type(scheme).__name__)


@apply_operator.register(FiniteVolumeScheme)
def _apply_operator_burgers(scheme: FiniteVolumeScheme, grid: Grid, bc: Boundary,
                            t: Any, u: torch.Tensor) -> torch.Tensor:
    """``schemes.py:339-346`` fused: periodic ghost fill, WENO-JS5 faces and
    the conservative difference in one kernel."""
    from pyshocks_amd.scalar import PeriodicBoundary, _eps_of, _rec_of, rusanov_nu

    if not isinstance(bc, PeriodicBoundary):
        return _apply_operator_ghosts(scheme, grid, bc, t, u)
    nu = rusanov_nu(grid, scheme.alpha) if isinstance(scheme, Rusanov) else None
    _lib.require_device(u, nu)
    out = torch.empty_like(u)
    from pyshocks_amd.reconstruction import MUSCL, LearnedWENOJS53

    # Lax-Friedrichs: the global speed of the ghost-filled state, whose
    # ghosts are images of interior cells
    speed = max_abs_interior(u, grid.nghosts) if isinstance(scheme, LaxFriedrichs) \
        else None
    if isinstance(scheme.rec, LearnedWENOJS53) and flux_kind(scheme) not in (
            _lib.PSA_FLUX_RUSANOV, _lib.PSA_FLUX_BURGERS_UPWIND,
            _lib.PSA_FLUX_LAX_FRIEDRICHS, _lib.PSA_FLUX_ENGQUIST_OSHER):
        raise NotImplementedError(
            "learned weights are implemented for the Rusanov, Godunov, "
            "Lax-Friedrichs and Engquist-Osher fluxes")
    if isinstance(scheme.rec, (LearnedWENOJS53, MUSCL)):
        # the operator stage of the descriptor carries the learned weights or
        # the limiter
        from pyshocks_amd.timestepping import BoundOperator, _scheme_plan

        sp = _scheme_plan(BoundOperator(scheme, grid, bc))
        assert sp is not None
        desc, _keep = sp
        if speed is not None:
            desc.velocity = _lib.ptr(speed)
        _lib.call("psa_ssprk33_stage", ctypes.byref(desc), 0, None, _lib.ptr(u),
                  _lib.ptr(out), None, None, _lib.stream_of(u))
        return out
    _lib.call("psa_apply_operator_rec", _lib.dtype_code(u.dtype), flux_kind(scheme),
              _rec_of(scheme), _lib.ptr(u), _lib.ptr(speed), _lib.ptr(nu),
              _lib.ptr(out), u.numel(), grid.nghosts, grid.h, _eps_of(scheme),
              _lib.stream_of(u))
    return out


def _apply_operator_ghosts(scheme: Any, grid: Grid, bc: Boundary, t: Any,
                           u: torch.Tensor) -> torch.Tensor:
    """``schemes.py:339-346`` with a non-periodic boundary: the ghost fill of
    ``bc`` (Dirichlet, Neumann), then the operator stage on the array as it
    is (``bc = PSA_BC_GIVEN``); the Lax-Friedrichs speed is the maximum over
    the whole filled array (``scalar.py:276``)."""
    from pyshocks_amd.schemes import apply_boundary
    from pyshocks_amd.timestepping import _scheme_desc

    sd = _scheme_desc(scheme, grid)
    if sd is None:
        raise NotImplementedError(f"{type(scheme).__name__} is not on the GPU path")
    desc, _keep = sd
    desc.bc = _lib.PSA_BC_GIVEN
    # the velocity array is reconstructed as the reference does, ends included
    desc.velocity_is_const = 0
    ub = apply_boundary(bc, grid, t, u)
    speed = None
    if desc.flux == _lib.PSA_FLUX_LAX_FRIEDRICHS:
        speed = max_abs_interior(ub, 0)
        desc.velocity = _lib.ptr(speed)
    out = torch.empty_like(u)
    _lib.call("psa_ssprk33_stage", ctypes.byref(desc), 0, None, _lib.ptr(ub),
              _lib.ptr(out), None, None, _lib.stream_of(u))
    return out


# {{{ ESWENO32


@dataclass(frozen=True)
class ESWENO32(FiniteVolumeScheme):
    """Energy-stable WENO (``burgers/schemes.py:202-256``): the upwind flux of
    the ES-WENO3 reconstruction plus the dissipative flux of its first
    weight; :func:`~pyshocks_amd.schemes.bind` sets the reconstruction's
    ``eps`` and ``delta`` from the grid. ``make_ssprk33`` runs it through
    the fused stage and whole-step kernels (``PSA_FLUX_ESWENO`` with the
    descriptor's ``es_delta``); ``numerical_flux`` and ``apply_operator``
    use the standalone kernels (upwind faces, dissipation, difference)."""

    def __post_init__(self) -> None:
        from pyshocks_amd.reconstruction import ESWENO32 as ESWENO32Rec

        if not isinstance(self.rec, ESWENO32Rec):
            raise TypeError("ESWENO32 scheme requires the ESWENO32 reconstruction.")


@bind.register(ESWENO32)
def _bind_burgers_esweno32(scheme: ESWENO32, grid: Grid, bc: Boundary) -> ESWENO32:
    """``burgers/schemes.py:219-227``: ``es_weno_parameters`` of a unit state."""
    from pyshocks_amd.weno import es_weno_parameters

    eps, delta = es_weno_parameters(grid, torch.ones_like(grid.x))
    return replace(scheme, rec=replace(scheme.rec, eps=float(eps), delta=float(delta)))


def esweno_dissipation(rec: Any, u: torch.Tensor, fnum: torch.Tensor,
                       weight: int) -> torch.Tensor:
    """``fnum`` plus the dissipative flux of the ES-WENO weight ``weight``
    of ``u`` (``burgers/schemes.py:240-254``), in place."""
    _lib.require_device(u, fnum)
    _lib.call("psa_esweno_dissipation", _lib.dtype_code(u.dtype), _lib.ptr(u),
              _lib.ptr(fnum), u.numel(), float(rec.eps), float(rec.delta), weight,
              _lib.stream_of(u))
    return fnum


@numerical_flux.register(ESWENO32)
def _numerical_flux_burgers_esweno32(scheme: ESWENO32, grid: Grid, bc: Boundary,
                                     t: Any, u: torch.Tensor) -> torch.Tensor:
    from pyshocks_amd.scalar import _eps_of, _numerical_flux

    _lib.require_device(u)
    fnum = _numerical_flux(_lib.PSA_FLUX_BURGERS_UPWIND, u, None, None, _eps_of(scheme),
                           _lib.PSA_REC_ESWENO32)
    return esweno_dissipation(scheme.rec, u, fnum, 0)


@apply_operator.register(ESWENO32)
def _apply_operator_burgers_esweno32(scheme: ESWENO32, grid: Grid, bc: Boundary,
                                     t: Any, u: torch.Tensor) -> torch.Tensor:
    """``schemes.py:466-473``: boundary, face fluxes, their difference."""
    from pyshocks_amd.schemes import apply_boundary

    ub = apply_boundary(bc, grid, t, u)
    fnum = numerical_flux(scheme, grid, bc, t, ub)
    out = torch.empty_like(u)
    _lib.call("psa_flux_divergence", _lib.dtype_code(u.dtype), _lib.ptr(fnum), None,
              _lib.ptr(out), u.numel(), grid.h, _lib.stream_of(u))
    return out


# }}}


# {{{ registry

_SCHEMES: dict[str, type[BurgersScheme]] = {
    "default": LaxFriedrichs,
    "godunov": Godunov,
    "rusanov": Rusanov,
    "lf": LaxFriedrichs,
    "eo": EngquistOsher,
    "esweno32": ESWENO32,
}

_NOT_ON_GPU = ("ssmuscl", "ssweno242", "splitrusanov")


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
