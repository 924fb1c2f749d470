"""Scalar numerical fluxes and the periodic boundary of the hot path.

Mirrors ``scalar.py`` of pyshocks: ``scalar_flux_upwind`` (``:91-132``),
``scalar_flux_rusanov`` (``:192-249``), ``scalar_flux_lax_friedrichs``
(``:258-278``), ``scalar_flux_engquist_osher`` (``:287-322``) and
``PeriodicBoundary`` /
``apply_boundary`` / ``evaluate_boundary`` (``:521-547``). The fluxes run in
the library's face kernel, which reconstructs both states of every face with
WENO-JS5 and evaluates the Burgers flux ``u^2 / 2`` inline, so they accept
Burgers schemes with a WENO-JS5 reconstruction.
"""

from __future__ import annotations

import ctypes

from dataclasses import dataclass
from typing import TYPE_CHECKING, Any

import torch

from pyshocks_amd import _lib
from pyshocks_amd.schemes import (
    Boundary,
    BoundaryType,
    ConservationLawScheme,
    apply_boundary,
    evaluate_boundary,
)

if TYPE_CHECKING:
    from pyshocks_amd.grid import Grid


def _eps_of(scheme: ConservationLawScheme) -> float:
    _rec_of(scheme)
    return float(getattr(scheme.rec, "eps", 0.0))


def _rec_of(scheme: ConservationLawScheme) -> int:
    """The ``PSA_REC_*`` code of the scheme's reconstruction (WENO-JS5, the
    learned variant of it, WENO-JS3 or ES-WENO3)."""
    from pyshocks_amd.reconstruction import LearnedWENOJS53, rec_code

    if isinstance(scheme.rec, LearnedWENOJS53):
        return _lib.PSA_REC_WENOJS53
    return rec_code(scheme.rec)


def _check_burgers(scheme: ConservationLawScheme) -> None:
    from pyshocks_amd.burgers import BurgersScheme

    if not isinstance(scheme, BurgersScheme):
        raise NotImplementedError(
            f"the GPU flux kernels evaluate the Burgers flux; got {type(scheme).__name__}")


def rusanov_nu(grid: Grid, alpha: float) -> torch.Tensor | None:
    """Artificial viscosity ``df ** (alpha - 1)`` (``scalar.py:230-234``), or
    ``None`` for ``|alpha - 1| <= 1e-8``."""
    if abs(alpha - 1.0) > 1.0e-8:
        return (grid.df ** (alpha - 1)).contiguous()
    return None


def _learned(scheme: ConservationLawScheme) -> Any:
    from pyshocks_amd.reconstruction import LearnedWENOJS53

    return scheme.rec if isinstance(scheme.rec, LearnedWENOJS53) else None


def _learned_flux(rec: Any, kind: int, u: torch.Tensor, nu: torch.Tensor | None,
                  speed: torch.Tensor | None = None) -> torch.Tensor:
    """Face fluxes of the learned reconstruction (``scalar.py:125-130``,
    ``:238-248``, ``:276-277`` and ``:309-319`` on the learned states);
    ``speed`` is the Lax-Friedrichs device scalar max|a|."""
    from pyshocks_amd.learned import learned_reconstruct

    ul, ur = learned_reconstruct(rec.weights, u, rec.eps)
    fl, fr = ul * ul / 2, ur * ur / 2
    if kind in (_lib.PSA_FLUX_RUSANOV, _lib.PSA_FLUX_LAX_FRIEDRICHS):
        if kind == _lib.PSA_FLUX_RUSANOV:
            a = torch.maximum(u[1:].abs(), u[:-1].abs())
        else:
            assert speed is not None
            a = speed
        nu_ = 1.0 if nu is None else nu
        f = 0.5 * (fl[1:] + fr[:-1]) - 0.5 * a * nu_ * (ul[1:] - ur[:-1])
    elif kind == _lib.PSA_FLUX_ENGQUIST_OSHER:
        p = torch.clamp(ur[:-1], min=0.0)
        q = torch.clamp(ul[1:], max=0.0)
        f = p * p / 2 + q * q / 2
    else:
        f = torch.where((ur[:-1] + ul[1:]) / 2 > 0, fr[:-1], fl[1:])
    return torch.nn.functional.pad(f, (1, 1))


def scalar_flux_upwind(scheme: ConservationLawScheme, grid: Grid, bc: BoundaryType,
                       t: Any, a: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
    """Upwind flux on the averaged reconstructed speed; the speed must be the
    state itself (Burgers, ``burgers/schemes.py:84-89``)."""
    assert u.shape[0] == grid.x.numel()
    assert a.shape == u.shape
    _check_burgers(scheme)
    if a is not u and not torch.equal(a, u):
        raise NotImplementedError("the GPU upwind flux takes the speed a = u")
    rec = _learned(scheme)
    if rec is not None:
        return _learned_flux(rec, _lib.PSA_FLUX_BURGERS_UPWIND, u, None)
    return _numerical_flux(_lib.PSA_FLUX_BURGERS_UPWIND, u, None, None,
                           _eps_of(scheme), _rec_of(scheme), scheme=scheme, grid=grid)


def scalar_flux_rusanov(scheme: ConservationLawScheme, grid: Grid, bc: BoundaryType,
                        t: Any, a: torch.Tensor, u: torch.Tensor,
                        alpha: float = 1.0) -> torch.Tensor:
    """Rusanov flux with the local speed ``max(|a_j|, |a_{j+1}|)`` of the cell
    values; the speed must be the state itself (Burgers,
    ``burgers/schemes.py:111-118``)."""
    assert u.shape[0] == grid.x.numel()
    _check_burgers(scheme)
    if a is not u and not torch.equal(a, u):
        raise NotImplementedError("the GPU Rusanov flux takes the speed a = u")
    rec = _learned(scheme)
    if rec is not None:
        return _learned_flux(rec, _lib.PSA_FLUX_RUSANOV, u, rusanov_nu(grid, alpha))
    return _numerical_flux(_lib.PSA_FLUX_RUSANOV, u, None, rusanov_nu(grid, alpha),
                           _eps_of(scheme), _rec_of(scheme), scheme=scheme, grid=grid)


def scalar_flux_lax_friedrichs(scheme: ConservationLawScheme, grid: Grid,
                               bc: BoundaryType, t: Any, a: torch.Tensor,
                               u: torch.Tensor, alpha: float = 1.0) -> torch.Tensor:
    """Rusanov flux with the global speed ``max|a|`` over the whole array
    (``scalar.py:276-277``); ``a`` may be any array of ``u``'s shape."""
    from pyshocks_amd.reductions import max_abs_interior

    assert u.shape[0] == grid.x.numel()
    _check_burgers(scheme)
    _lib.require_device(a)
    if a.shape != u.shape or a.dtype != u.dtype:
        raise ValueError("the speed array must have the shape and dtype of u")
    amax = max_abs_interior(a.contiguous(), 0)
    rec = _learned(scheme)
    if rec is not None:
        return _learned_flux(rec, _lib.PSA_FLUX_LAX_FRIEDRICHS, u, rusanov_nu(grid, alpha),
                             amax)
    return _numerical_flux(_lib.PSA_FLUX_LAX_FRIEDRICHS, u, amax,
                           rusanov_nu(grid, alpha), _eps_of(scheme), _rec_of(scheme),
                           scheme=scheme, grid=grid)


def scalar_flux_engquist_osher(scheme: ConservationLawScheme, grid: Grid,
                               bc: BoundaryType, t: Any, a: torch.Tensor,
                               u: torch.Tensor, omega: float = 0.0) -> torch.Tensor:
    """``f(max(u_l, omega)) + f(min(u_r, omega)) - f(omega)`` on the WENO-JS
    face states (``scalar.py:309-319``); the Burgers flux has its minimum at
    ``omega = 0``, the only value the kernel takes. ``a`` is unused, as in the
    reference's WENO reconstruction."""
    assert u.shape[0] == grid.x.numel()
    _check_burgers(scheme)
    if omega != 0.0:
        raise NotImplementedError("the GPU Engquist-Osher flux takes omega = 0")
    rec = _learned(scheme)
    if rec is not None:
        return _learned_flux(rec, _lib.PSA_FLUX_ENGQUIST_OSHER, u, None)
    return _numerical_flux(_lib.PSA_FLUX_ENGQUIST_OSHER, u, None, None,
                           _eps_of(scheme), _rec_of(scheme), scheme=scheme, grid=grid)


def _numerical_flux(kind: int, u: torch.Tensor, velocity: torch.Tensor | None,
                    nu: torch.Tensor | None, eps: float,
                    rec: int = _lib.PSA_REC_WENOJS53, *, scheme: Any = None,
                    grid: Any = None) -> torch.Tensor:
    """Face fluxes of u as given; ``velocity`` is the advection velocity or
    the Lax-Friedrichs speed. MUSCL (its limiter in the descriptor) needs the
    scheme and the grid."""
    _lib.require_device(u, velocity, nu)
    fnum = torch.empty(u.numel() + 1, dtype=u.dtype, device=u.device)
    if rec in (_lib.PSA_REC_MUSCL, _lib.PSA_REC_MUSCLS):
        from pyshocks_amd.limiters import limiter_code

        assert scheme is not None and grid is not None
        desc = _lib.PsaSolverDesc()
        desc.dtype = _lib.dtype_code(u.dtype)
        desc.flux = kind
        desc.g = grid.nghosts
        desc.n = u.numel() - 2 * grid.nghosts
        desc.dx = grid.h
        desc.rec = rec
        desc.limiter, desc.limiter_theta = limiter_code(scheme.rec.lm)
        desc.muscl_atol = float(scheme.rec.atol)
        desc.nu = _lib.ptr(nu)
        speed = velocity if kind == _lib.PSA_FLUX_LAX_FRIEDRICHS else None
        if kind == _lib.PSA_FLUX_ADVECTION_UPWIND:
            desc.velocity = _lib.ptr(velocity)
        _lib.call("psa_numerical_flux_desc", ctypes.byref(desc), _lib.ptr(u),
                  _lib.ptr(speed), _lib.ptr(fnum), _lib.stream_of(u))
        return fnum
    _lib.call("psa_numerical_flux_rec", _lib.dtype_code(u.dtype), kind, rec,
              _lib.ptr(u), _lib.ptr(velocity), _lib.ptr(nu), _lib.ptr(fnum),
              u.numel(), eps, _lib.stream_of(u))
    return fnum


# {{{ Dirichlet and Neumann


@dataclass(frozen=True)
class OneSidedBoundary(Boundary):
    """A boundary condition on one side: ``side`` is ``-1`` (left) or
    ``+1`` (right) (``scalar.py:333-342``)."""

    side: int


@dataclass(frozen=True)
class TwoSidedBoundary(Boundary):
    """One boundary condition on each side (``scalar.py:345-382``); each side
    is applied in turn."""

    left: OneSidedBoundary
    right: OneSidedBoundary

    def __post_init__(self) -> None:
        assert isinstance(self.left, OneSidedBoundary)
        assert isinstance(self.right, OneSidedBoundary)
        if self.left.side != -1:
            raise ValueError("Left boundary has incorrect side.")
        if self.right.side != +1:
            raise ValueError("Right boundary has incorrect side.")

    @property
    def boundary_type(self) -> BoundaryType:
        if self.left.boundary_type != self.right.boundary_type:
            raise NotImplementedError("Different boundaries on each side.")
        return self.left.boundary_type


@apply_boundary.register(TwoSidedBoundary)
def _apply_boundary_two_sided(bc: TwoSidedBoundary, grid: Grid, t: Any,
                              u: torch.Tensor) -> torch.Tensor:
    u = apply_boundary(bc.left, grid, t, u)
    return apply_boundary(bc.right, grid, t, u)


@evaluate_boundary.register(TwoSidedBoundary)
def _evaluate_boundary_two_sided(bc: TwoSidedBoundary, grid: Grid, t: Any,
                                 u: torch.Tensor) -> torch.Tensor:
    return (evaluate_boundary(bc.left, grid, t, u)
            + evaluate_boundary(bc.right, grid, t, u))


@dataclass(frozen=True)
class DirichletBoundary(OneSidedBoundary):
    """``u(t, a) = g(t, x)`` in the ghost cells of ``side``; ``g`` is evaluated
    on the ghost cell centres (a device tensor) and may return a scalar. A
    ``g`` that is not callable (a number, or the ghost values) is the same
    value at every time: the fused steps then fill the ghost cells on the
    device."""

    g: Any

    @property
    def boundary_type(self) -> BoundaryType:
        return BoundaryType.Dirichlet


@apply_boundary.register(DirichletBoundary)
def _apply_boundary_scalar_dirichlet(bc: DirichletBoundary, grid: Grid, t: Any,
                                     u: torch.Tensor) -> torch.Tensor:
    """``scalar.py:441-449``: the ghost cells ``grid.g_[side]`` get
    ``g(t, x[ghosts])``."""
    assert u.numel() == grid.x.numel()
    _lib.require_device(u)
    g = grid.nghosts
    xg = grid.x[:g] if bc.side == -1 else grid.x[grid.n - g:]
    vals = dirichlet_values(bc, t, xg, u.dtype)
    out = torch.empty_like(u)
    left, right = (vals, None) if bc.side == -1 else (None, vals)
    _lib.call("psa_apply_boundary_dirichlet", _lib.dtype_code(u.dtype), _lib.ptr(u),
              _lib.ptr(out), u.numel(), g, _lib.ptr(left), _lib.ptr(right),
              _lib.stream_of(u))
    return out


def dirichlet_values(bc: DirichletBoundary, t: Any, xg: torch.Tensor,
                     dtype: torch.dtype) -> torch.Tensor:
    """The ``nghosts`` ghost values of a Dirichlet side at time ``t`` (``xg``:
    its ghost cell centres)."""
    vals = bc.g(t, xg) if callable(bc.g) else bc.g
    vals = torch.as_tensor(vals, dtype=dtype, device=xg.device)
    return torch.broadcast_to(vals, xg.shape).contiguous()


def neumann_gradient(bc: NeumannBoundary, t: Any) -> float:
    """``g(t)`` of a Neumann side (the number itself for a constant)."""
    return float(bc.g(t)) if callable(bc.g) else float(bc.g)


def make_dirichlet_boundary(ga: Any, gb: Any = None) -> TwoSidedBoundary:
    """``scalar.py:452-460``."""
    if gb is None:
        gb = ga
    return TwoSidedBoundary(left=DirichletBoundary(side=-1, g=ga),
                            right=DirichletBoundary(side=+1, g=gb))


@dataclass(frozen=True)
class NeumannBoundary(OneSidedBoundary):
    """``side du/dx (t, a) = g(t)`` by a mirror of the interior cells
    (``scalar.py:466-511``); ``g`` returns a number. A ``g`` that is a number
    is a constant gradient: the fused steps then fill the ghost cells on the
    device."""

    g: Any

    @property
    def boundary_type(self) -> BoundaryType:
        return BoundaryType.Neumann


@apply_boundary.register(NeumannBoundary)
def _apply_boundary_scalar_neumann(bc: NeumannBoundary, grid: Grid, t: Any,
                                   u: torch.Tensor) -> torch.Tensor:
    """``scalar.py:477-500``: ``u[to] = u[from] + side (x[from] - x[to]) g(t)``
    with the first (last) ``nghosts`` interior cells mirrored into the ghost
    layer, on a uniform grid ``x[from] - x[to] = (from - to) dx``."""
    assert u.numel() == grid.x.numel()
    _lib.require_device(u)
    gn = neumann_gradient(bc, t)
    out = torch.empty_like(u)
    left = bc.side == -1
    _lib.call("psa_apply_boundary_neumann", _lib.dtype_code(u.dtype), _lib.ptr(u),
              _lib.ptr(out), u.numel(), grid.nghosts, grid.h,
              gn if left else 0.0, 0.0 if left else gn, 1 if left else 2,
              _lib.stream_of(u))
    return out


def make_neumann_boundary(ga: Any, gb: Any = None) -> TwoSidedBoundary:
    """``scalar.py:503-511``."""
    if gb is None:
        gb = ga
    return TwoSidedBoundary(left=NeumannBoundary(side=-1, g=ga),
                            right=NeumannBoundary(side=+1, g=gb))


# }}}


# {{{ periodic


@dataclass(frozen=True)
class PeriodicBoundary(Boundary):
    """Periodic boundary conditions for one-dimensional domains."""

    @property
    def boundary_type(self) -> BoundaryType:
        return BoundaryType.Periodic


@apply_boundary.register(PeriodicBoundary)
def _apply_boundary_scalar_periodic(bc: PeriodicBoundary, grid: Grid, t: Any,
                                    u: torch.Tensor) -> torch.Tensor:
    assert u.numel() == grid.x.numel()
    _lib.require_device(u)
    out = torch.empty_like(u)
    _lib.call("psa_apply_boundary_periodic", _lib.dtype_code(u.dtype), _lib.ptr(u),
              _lib.ptr(out), u.numel(), grid.nghosts, _lib.stream_of(u))
    return out


@evaluate_boundary.register(PeriodicBoundary)
def _evaluate_boundary_scalar_periodic(bc: PeriodicBoundary, grid: Grid, t: Any,
                                       u: torch.Tensor) -> torch.Tensor:
    return torch.zeros_like(u)


# }}}
