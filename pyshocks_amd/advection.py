"""Linear advection ``u_t + a u_x = 0`` with the Godunov (upwind) scheme.

Mirrors ``advection/schemes.py`` of pyshocks (``AdvectionScheme``
``:41-48``, ``predict_timestep`` ``:51-59``, ``apply_operator`` ``:62-73``,
``upwind_flux`` ``:100-114``, ``Godunov`` ``:118-129``) and the registry of
``advection/__init__.py:59-96``.
"""

from __future__ import annotations

from dataclasses import dataclass, fields
from typing import TYPE_CHECKING, Any

import torch

from pyshocks_amd import _lib
from pyshocks_amd.reductions import max_abs_interior
from pyshocks_amd.schemes import (
    Boundary,
    FiniteVolumeSchemeBase,
    SchemeBase,
    apply_operator,
    numerical_flux,
    predict_timestep,
)

if TYPE_CHECKING:
    from pyshocks_amd.grid import Grid


@dataclass(frozen=True)
class AdvectionScheme(SchemeBase):
    """Base class for numerical schemes for linear advection."""

    velocity: torch.Tensor | None
    """Advection velocity at the cell centres (shape of ``grid.x``)."""


@predict_timestep.register(AdvectionScheme)
def _predict_timestep_advection(scheme: AdvectionScheme, grid: Grid, bc: Boundary,
                                t: Any, u: torch.Tensor) -> torch.Tensor:
    assert scheme.velocity is not None
    amax = max_abs_interior(scheme.velocity, grid.nghosts)
    return grid.dx_min / amax


@dataclass(frozen=True)
class FiniteVolumeScheme(AdvectionScheme, FiniteVolumeSchemeBase):
    """Finite volume schemes for linear advection."""


def _eps_and_check(scheme: AdvectionScheme) -> float:
    _rec_and_check(scheme)
    return float(getattr(scheme.rec, "eps", 0.0))


def _rec_and_check(scheme: AdvectionScheme) -> int:
    """``PSA_REC_*`` of the reconstruction (learned weights are for Burgers)."""
    from pyshocks_amd.reconstruction import rec_code

    return rec_code(scheme.rec)


def upwind_flux(scheme: AdvectionScheme, grid: Grid, bc: Boundary,
                u: torch.Tensor) -> torch.Tensor:
    """``F = ur if (ar + al) / 2 > 0 else ul`` with ``u`` and the velocity
    both reconstructed by the scheme's reconstruction."""
    assert scheme.velocity is not None
    assert scheme.rec is not None
    assert u.shape[0] == grid.x.numel()

    from pyshocks_amd.scalar import _numerical_flux

    return _numerical_flux(_lib.PSA_FLUX_ADVECTION_UPWIND, u, scheme.velocity,
                           None, _eps_and_check(scheme), _rec_and_check(scheme),
                           scheme=scheme, grid=grid)


@dataclass(frozen=True)
class Godunov(FiniteVolumeScheme):
    """A Godunov (upwind) scheme for the advection equation."""


@numerical_flux.register(Godunov)
def _numerical_flux_advection_godunov(scheme: Godunov, grid: Grid, bc: Boundary,
                                      t: Any, u: torch.Tensor) -> torch.Tensor:
    return upwind_flux(scheme, grid, bc, u)


@apply_operator.register(Godunov)
def _apply_operator_advection(scheme: Godunov, grid: Grid, bc: Boundary, t: Any,
                              u: torch.Tensor) -> torch.Tensor:
    """``-velocity * (F[1:] - F[:-1]) / dx`` with the periodic ghost fill, in
    one kernel."""
    from pyshocks_amd.scalar import PeriodicBoundary

    assert scheme.velocity is not None
    if not isinstance(bc, PeriodicBoundary):
        from pyshocks_amd.burgers import _apply_operator_ghosts

        _lib.require_device(u, scheme.velocity)
        return _apply_operator_ghosts(scheme, grid, bc, t, u)
    _lib.require_device(u, scheme.velocity)
    out = torch.empty_like(u)
    if _rec_and_check(scheme) in (_lib.PSA_REC_MUSCL, _lib.PSA_REC_MUSCLS):
        # the limiter travels in the descriptor: the operator stage
        import ctypes

        from pyshocks_amd.timestepping import BoundOperator, _scheme_plan

        sp = _scheme_plan(BoundOperator(scheme, grid, bc))
        assert sp is not None
        desc, _keep = sp
        desc.velocity_is_const = 0
        _lib.call("psa_ssprk33_stage", ctypes.byref(desc), 0, None, _lib.ptr(u),
                  _lib.ptr(out), None, None, _lib.stream_of(u))
        return out
    _lib.call("psa_apply_operator_rec", _lib.dtype_code(u.dtype),
              _lib.PSA_FLUX_ADVECTION_UPWIND, _rec_and_check(scheme), _lib.ptr(u),
              _lib.ptr(scheme.velocity), None, _lib.ptr(out), u.numel(), grid.nghosts,
              grid.h, _eps_and_check(scheme), _lib.stream_of(u))
    return out


Upwind = Godunov


@dataclass(frozen=True)
class ESWENO32(FiniteVolumeScheme):
    """Energy-stable WENO for advection (``advection/schemes.py:135-186``):
    a Lax-Friedrichs flux of the reconstructed ``u`` and ``f = velocity u``
    with the global speed ``max|velocity|``, plus the dissipative flux of the
    second ES-WENO weight of ``u``. The GPU path takes the ES-WENO3
    reconstruction only (the reference also accepts others, without the
    dissipative term); it runs through the operator kernels and the generic
    step loop."""


@numerical_flux.register(ESWENO32)
def _numerical_flux_advection_esweno32(scheme: ESWENO32, grid: Grid, bc: Boundary,
                                       t: Any, u: torch.Tensor) -> torch.Tensor:
    from pyshocks_amd.reconstruction import ESWENO32 as ESWENO32Rec
    from pyshocks_amd.reductions import max_abs_interior

    assert scheme.velocity is not None
    if not isinstance(scheme.rec, ESWENO32Rec):
        raise NotImplementedError(
            "the GPU ES-WENO advection flux takes the ES-WENO3 reconstruction")
    _lib.require_device(u, scheme.velocity)
    amax = max_abs_interior(scheme.velocity, 0)
    fnum = torch.empty(u.numel() + 1, dtype=u.dtype, device=u.device)
    _lib.call("psa_esweno_advection_flux", _lib.dtype_code(u.dtype), _lib.ptr(u),
              _lib.ptr(scheme.velocity), _lib.ptr(amax), _lib.ptr(fnum), u.numel(),
              float(scheme.rec.eps), float(scheme.rec.delta), _lib.stream_of(u))
    return fnum


@apply_operator.register(ESWENO32)
def _apply_operator_advection_esweno32(scheme: ESWENO32, grid: Grid, bc: Boundary,
                                       t: Any, u: torch.Tensor) -> torch.Tensor:
    """``advection/schemes.py:62-73``: boundary, face fluxes and
    ``-velocity (F[1:] - F[:-1]) / dx``."""
    from pyshocks_amd.schemes import apply_boundary

    ub = apply_boundary(bc, grid, t, u)
    fnum = numerical_flux(scheme, grid, bc, t, ub)
    out = torch.empty_like(u)
    _lib.call("psa_flux_divergence", _lib.dtype_code(u.dtype), _lib.ptr(fnum),
              _lib.ptr(scheme.velocity), _lib.ptr(out), u.numel(), grid.h,
              _lib.stream_of(u))
    return out

# {{{ registry

_SCHEMES: dict[str, type[AdvectionScheme]] = {
    "default": Godunov,
    "godunov": Godunov,
    "upwind": Godunov,
    "esweno32": ESWENO32,
}

_NOT_ON_GPU = ("sbp", "splitgodunov")


def scheme_ids() -> tuple[str, ...]:
    return tuple(_SCHEMES.keys())


def make_scheme_from_name(name: str, **kwargs: Any) -> AdvectionScheme:
    """``advection/__init__.py:76-96``: ``velocity`` defaults to ``None``."""
    cls = _SCHEMES.get(name)
    if cls is None:
        if name in _NOT_ON_GPU:
            raise NotImplementedError(
                f"Scheme {name!r} is not part of the GPU hot path; available: "
                f"{', '.join(scheme_ids())}.")
        raise ValueError(f"Scheme {name!r} not found. Try one of "
                         f"{', '.join(scheme_ids())}.")

    if "velocity" not in kwargs:
        kwargs["velocity"] = None

    return cls(**{f.name: kwargs[f.name] for f in fields(cls) if f.name in kwargs})


# }}}
