"""Reconstruction interface and WENO-JS5.

Mirrors ``reconstruction.py`` of pyshocks: ``Reconstruction`` (``:49-76``),
``reconstruct`` (``:80-118``), ``WENOJS`` (``:306-316``), ``WENOJS32``
(``:320-333``), ``WENOJS53`` (``:336-348``), the WENO-JS registration
(``:358-377``), ``ESWENO32`` and its registration (``:387-434``), ``MUSCL``
(``:172-252``, the flux-limited form) and ``make_reconstruction_from_name``
(``:528-564``).
"""

from __future__ import annotations

from dataclasses import dataclass, field, fields
from functools import singledispatch
from typing import TYPE_CHECKING, Any, ClassVar

import torch

from pyshocks_amd import _lib, weno

if TYPE_CHECKING:
    from pyshocks_amd.grid import Grid
    from pyshocks_amd.schemes import BoundaryType


@dataclass(frozen=True)
class Reconstruction:
    """A reconstruction algorithm."""

    @property
    def name(self) -> str:
        return type(self).__name__.lower()

    @property
    def order(self) -> int:
        raise NotImplementedError(type(self).__name__)

    @property
    def stencil_width(self) -> int:
        raise NotImplementedError(type(self).__name__)


@singledispatch
def reconstruct(rec: Reconstruction, grid: Grid, bc: BoundaryType, f: torch.Tensor,
                u: torch.Tensor, wavespeed: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """``(fl, fr)``: ``fr[i]`` at ``x_{i+1/2}``, ``fl[i]`` at ``x_{i-1/2}``."""
    raise NotImplementedError(type(rec).__name__)


@dataclass(frozen=True)
class WENOJS(Reconstruction):
    """A WENO-JS reconstruction."""

    eps: float
    s: ClassVar[weno.Stencil]

    @property
    def stencil_width(self) -> int:
        return self.order


@dataclass(frozen=True)
class WENOJS53(WENOJS):
    """Fifth-order WENO-JS (third order at discontinuities)."""

    eps: float = 1.0e-12

    def __post_init__(self) -> None:
        object.__setattr__(self, "s", weno.weno_js_53_coefficients())

    @property
    def order(self) -> int:
        return 3


@dataclass(frozen=True)
class WENOJS32(WENOJS):
    """Third-order WENO-JS (second order at discontinuities)."""

    eps: float = 1.0e-6

    def __post_init__(self) -> None:
        object.__setattr__(self, "s", weno.weno_js_32_coefficients())

    @property
    def order(self) -> int:
        return 2


@dataclass(frozen=True)
class ESWENO32(Reconstruction):
    """The WENO-JS3 stencils with the ES-WENO weights
    ``d (1 + tau / (eps + beta))`` (:func:`~pyshocks_amd.weno.es_weno_weights`).
    ``delta`` belongs to the ES-WENO scheme's dissipation, not to the
    weights."""

    eps: float = 1.0e-6
    delta: float = 1.0e-6
    s: ClassVar[weno.Stencil]

    def __post_init__(self) -> None:
        object.__setattr__(self, "s", weno.weno_js_32_coefficients())

    @property
    def order(self) -> int:
        return 2

    @property
    def stencil_width(self) -> int:
        return 2


@dataclass(frozen=True)
class MUSCL(Reconstruction):
    """Second-order MUSCL with a flux limiter:
    ``ur = f + phi(r) / 2 (f[i+1] - f[i])``, ``ul = f - phi(1/r) / 2
    (f[i] - f[i-1])``, ``r`` the slope ratio cut off below ``atol``."""

    lm: Any
    atol: float = 1.0e-12

    @property
    def name(self) -> str:
        return f"{type(self).__name__}_{type(self.lm).__name__[:-7]}".lower()

    @property
    def order(self) -> int:
        return 2

    @property
    def stencil_width(self) -> int:
        return 2


@reconstruct.register(MUSCL)
def _reconstruct_muscl(rec: MUSCL, grid: Grid, bc: BoundaryType, f: torch.Tensor,
                       u: torch.Tensor, wavespeed: torch.Tensor,
                       ) -> tuple[torch.Tensor, torch.Tensor]:
    """The slope ratio is taken from ``f``, which is ``u`` at every call site
    of the reference's fluxes; a different ``u`` is not supported."""
    from pyshocks_amd.grid import UniformGrid
    from pyshocks_amd.limiters import limiter_code

    if not isinstance(grid, UniformGrid):
        raise NotImplementedError("MUSCL is only implemented for uniform grids.")
    assert grid.nghosts >= rec.stencil_width
    if u is not None and u is not f and not torch.equal(u, f):
        raise NotImplementedError("the GPU MUSCL takes the slope ratio of f itself")

    _lib.require_device(f)
    ul = torch.empty_like(f)
    ur = torch.empty_like(f)
    lm, theta = limiter_code(rec.lm)
    _lib.call("psa_muscl_reconstruct", _lib.dtype_code(f.dtype), lm, theta,
              float(rec.atol), _lib.ptr(f), _lib.ptr(ul), _lib.ptr(ur), f.numel(),
              _lib.stream_of(f))
    return ul, ur


@dataclass(frozen=True)
class MUSCLS(MUSCL):
    """Second-order MUSCL with a slope limiter (``reconstruction.py:255-296``):
    ``ur, ul = f +- dx / 2 du``, ``du`` the limited slope of ``f``
    (``limiters.py:101-372``), zero in the first and last cell of the array.
    The Koren limiter has no slope form."""


@reconstruct.register(MUSCLS)
def _reconstruct_muscls(rec: MUSCLS, grid: Grid, bc: BoundaryType, f: torch.Tensor,
                        u: torch.Tensor, wavespeed: torch.Tensor,
                        ) -> tuple[torch.Tensor, torch.Tensor]:
    from pyshocks_amd.grid import UniformGrid
    from pyshocks_amd.limiters import limiter_code

    if not isinstance(grid, UniformGrid):
        raise NotImplementedError("MUSCL is only implemented for uniform grids.")
    assert grid.nghosts >= rec.stencil_width

    _lib.require_device(f)
    ul = torch.empty_like(f)
    ur = torch.empty_like(f)
    lm, _theta = limiter_code(rec.lm)
    _lib.call("psa_muscls_reconstruct", _lib.dtype_code(f.dtype), lm,
              _lib.ptr(f), _lib.ptr(ul), _lib.ptr(ur), f.numel(), float(grid.h),
              _lib.stream_of(f))
    return ul, ur


def rec_code(rec: Reconstruction) -> int:
    """The library's ``PSA_REC_*`` code of a reconstruction on the GPU path."""
    if isinstance(rec, MUSCLS):
        return _lib.PSA_REC_MUSCLS
    if isinstance(rec, MUSCL):
        return _lib.PSA_REC_MUSCL
    if isinstance(rec, LearnedWENOJS53):
        raise NotImplementedError("learned weights have their own entry points")
    if isinstance(rec, WENOJS53):
        return _lib.PSA_REC_WENOJS53
    if isinstance(rec, WENOJS32):
        return _lib.PSA_REC_WENOJS32
    if isinstance(rec, ESWENO32):
        return _lib.PSA_REC_ESWENO32
    raise NotImplementedError(
        f"reconstruction {type(rec).__name__} is not on the GPU path")


def _reconstruct_weno(rec: Any, grid: Grid, f: torch.Tensor,
                      what: str) -> tuple[torch.Tensor, torch.Tensor]:
    from pyshocks_amd.grid import UniformGrid

    assert grid.nghosts >= rec.stencil_width

    if not isinstance(grid, UniformGrid):
        raise NotImplementedError(f"{what} is only implemented for uniform grids.")

    _lib.require_device(f)
    ul = torch.empty_like(f)
    ur = torch.empty_like(f)
    _lib.call("psa_reconstruct", _lib.dtype_code(f.dtype), rec_code(rec),
              _lib.ptr(f), _lib.ptr(ul), _lib.ptr(ur), f.numel(), float(rec.eps),
              _lib.stream_of(f))
    return ul, ur


@reconstruct.register(WENOJS)
def _reconstruct_wenojs(rec: WENOJS, grid: Grid, bc: BoundaryType,
                        f: torch.Tensor, u: torch.Tensor,
                        wavespeed: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    return _reconstruct_weno(rec, grid, f, "WENO-JS")


@reconstruct.register(ESWENO32)
def _reconstruct_esweno32(rec: ESWENO32, grid: Grid, bc: BoundaryType,
                          f: torch.Tensor, u: torch.Tensor,
                          wavespeed: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    return _reconstruct_weno(rec, grid, f, "ES-WENO")


@dataclass(frozen=True)
class LearnedWENOJS53(WENOJS53):
    """WENO-JS5 candidates with learned nonlinear weights (a small perceptron
    of the normalised stencil, :mod:`pyshocks_amd.learned`); cells flagged by
    the smoothness indicators keep the WENO-JS weights. No counterpart in the
    reference. ``weights`` defaults to the trained network shipped with the
    package."""

    weights: Any = field(default=None, compare=False, hash=False)

    def __post_init__(self) -> None:
        super().__post_init__()
        if self.weights is None:
            from pyshocks_amd.learned import MLPWeights

            object.__setattr__(self, "weights", MLPWeights.load())


@reconstruct.register(LearnedWENOJS53)
def _reconstruct_learned(rec: LearnedWENOJS53, grid: Grid, bc: BoundaryType,
                         f: torch.Tensor, u: torch.Tensor,
                         wavespeed: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    from pyshocks_amd.grid import UniformGrid
    from pyshocks_amd.learned import learned_reconstruct

    assert grid.nghosts >= rec.stencil_width
    if not isinstance(grid, UniformGrid):
        raise NotImplementedError("WENO-JS is only implemented for uniform grids.")
    return learned_reconstruct(rec.weights, f, rec.eps)


# {{{ registry

_RECONSTRUCTION: dict[str, type[Reconstruction]] = {
    "default": WENOJS53,
    "wenojs32": WENOJS32,
    "wenojs53": WENOJS53,
    "esweno32": ESWENO32,
    "muscl": MUSCL,
    "muscls": MUSCLS,
    "wenojs53-learned": LearnedWENOJS53,
}

# names the reference registers that are outside the GPU hot path
_NOT_ON_GPU = ("constant", "ssweno242")


def reconstruction_ids() -> tuple[str, ...]:
    return tuple(_RECONSTRUCTION.keys())


def make_reconstruction_from_name(name: str, **kwargs: Any) -> Reconstruction:
    """Unknown names raise ``ValueError`` (``reconstruction.py:553-560``);
    fields not in the class are ignored (``:562-564``)."""
    cls = _RECONSTRUCTION.get(name)
    if cls is None:
        if name in _NOT_ON_GPU:
            raise NotImplementedError(
                f"Reconstruction {name!r} is not part of the GPU hot path; "
                f"available: {', '.join(reconstruction_ids())}.")
        raise ValueError(
            f"Flux limiter {name!r} not found. Try one of "
            f"{', '.join(reconstruction_ids())}.")

    return cls(**{f.name: kwargs[f.name] for f in fields(cls) if f.name in kwargs})


# }}}
