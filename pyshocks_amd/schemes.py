"""Scheme base classes and the generic-function API of the hot path.

Mirrors ``schemes.py`` of pyshocks: ``SchemeBase`` (``:79-118``), ``bind``
(``:125-140``), ``apply_operator`` (``:143-167``), ``predict_timestep``
(``:170-191``), ``FiniteVolumeSchemeBase`` (``:194-199``),
``ConservationLawScheme`` (``:288-303``), ``flux`` (``:306-318``),
``numerical_flux`` (``:322-336``), the conservation-law operator
(``:339-346``), ``BoundaryType`` / ``Boundary`` / ``apply_boundary``
(``:394-443``). The generics are ``functools.singledispatch`` functions with
the reference's names, argument order and ``NotImplementedError`` for
unregistered types; registrations for the hot path call the HIP library.
"""

from __future__ import annotations

import enum
from abc import ABC, abstractmethod
from dataclasses import dataclass
from functools import singledispatch
from typing import TYPE_CHECKING, Any, TypeVar

import torch

if TYPE_CHECKING:
    from pyshocks_amd.grid import Grid
    from pyshocks_amd.reconstruction import Reconstruction


@dataclass(frozen=True)
class SchemeBase:
    """Numerical scheme for ``u_t = A(t, x, u, u_x)``."""

    rec: Reconstruction

    @property
    def name(self) -> str:
        return f"{type(self).__name__}_{self.rec.name}".lower()

    @property
    def order(self) -> int:
        return self.rec.order

    @property
    def stencil_width(self) -> int:
        return self.rec.stencil_width


SchemeT = TypeVar("SchemeT", bound=SchemeBase)


@singledispatch
def bind(scheme: SchemeT, grid: Grid, bc: Boundary) -> SchemeT:
    """Bind the scheme to a grid and boundary; identity for the hot-path
    schemes (``schemes.py:125-140``)."""
    if isinstance(scheme, SchemeBase):
        return scheme

    raise NotImplementedError(type(scheme).__name__)


@singledispatch
def apply_operator(scheme: SchemeBase, grid: Grid, bc: Boundary, t: Any,
                   u: torch.Tensor) -> torch.Tensor:
    """Right-hand side of the method of lines at cell centres."""
    raise NotImplementedError(type(scheme).__name__)


@singledispatch
def predict_timestep(scheme: SchemeBase, grid: Grid, bc: Boundary, t: Any,
                     u: torch.Tensor) -> torch.Tensor:
    """A priori time step estimate (to be multiplied by a Courant number)."""
    raise NotImplementedError(type(scheme).__name__)


@dataclass(frozen=True)
class FiniteVolumeSchemeBase(SchemeBase):
    """Finite volume-type scheme."""


@dataclass(frozen=True)
class FiniteDifferenceSchemeBase(SchemeBase):
    """Finite difference-type scheme (no hot-path registrations)."""


@dataclass(frozen=True)
class ConservationLawScheme(FiniteVolumeSchemeBase):
    """Scheme for ``u_t + f(u)_x = 0`` defined by :func:`flux` and
    :func:`numerical_flux`."""


@singledispatch
def flux(scheme: SchemeBase, t: Any, x: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
    """Physical flux."""
    raise NotImplementedError(type(scheme).__name__)


@singledispatch
def numerical_flux(scheme: SchemeBase, grid: Grid, bc: Boundary, t: Any,
                   u: torch.Tensor) -> torch.Tensor:
    """Numerical flux at every face (size of ``grid.f``)."""
    raise NotImplementedError(type(scheme).__name__)


# {{{ boundary conditions


@enum.unique
class BoundaryType(enum.Enum):
    Periodic = enum.auto()
    Dirichlet = enum.auto()
    Neumann = enum.auto()
    HomogeneousNeumann = enum.auto()


@dataclass(frozen=True)
class Boundary(ABC):
    """Boundary conditions for one-dimensional domains."""

    @property
    @abstractmethod
    def boundary_type(self) -> BoundaryType:
        """General form of the boundary condition."""


@singledispatch
def apply_boundary(bc: Boundary, grid: Grid, t: Any, u: torch.Tensor) -> torch.Tensor:
    """Copy of ``u`` with the ghost layer set by the boundary condition."""
    raise NotImplementedError(type(bc).__name__)


@singledispatch
def evaluate_boundary(bc: Boundary, grid: Grid, t: Any, u: torch.Tensor) -> torch.Tensor:
    """Boundary values at the domain boundaries, zero elsewhere."""
    raise NotImplementedError(type(bc).__name__)


# }}}
