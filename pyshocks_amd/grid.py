"""Uniform cell grids, cell averages and interior norms on the GPU.

Mirrors ``grid.py`` of pyshocks (``Grid`` ``:48-117``, ``UniformGrid``
``:121``, ``make_uniform_cell_grid`` ``:131-179``, ``Quadrature`` /
``make_leggauss_quadrature`` / ``cell_average`` ``:344-435``, ``norm`` /
``rnorm`` ``:445-531``). Arrays are ``torch.Tensor`` on the HIP device; the
layout is the reference's: ``n + 2 g`` cells with ``g`` ghost cells per side.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable

import numpy as np
import torch

from pyshocks_amd import _lib


def default_device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


@dataclass(frozen=True)
class Grid:
    a: float
    """Left domain bound for :math:`[a, b]`."""
    b: float
    """Right domain bound for :math:`[a, b]`."""
    nghosts: int
    """Number of ghost cells."""

    x: torch.Tensor
    """Cell centres, shape ``(n + 2 g,)``."""
    dx: torch.Tensor
    """Cell sizes."""
    f: torch.Tensor
    """Faces, shape ``(n + 2 g + 1,)``."""
    df: torch.Tensor
    """Distances between cell centres, shape ``(n + 2 g - 1,)``."""

    dx_min: torch.Tensor
    dx_max: torch.Tensor
    is_periodic: bool

    h: float
    """The uniform cell size as a Python float (``(b - a) / n``)."""

    @property
    def dtype(self) -> torch.dtype:
        return self.x.dtype

    @property
    def device(self) -> torch.device:
        return self.x.device

    @property
    def n(self) -> int:
        """Number of cells including ghosts (``grid.py:79-82``)."""
        return int(self.x.numel())

    @property
    def ncells(self) -> int:
        """Number of interior cells."""
        return self.n - 2 * self.nghosts

    @property
    def i_(self) -> slice:
        return np.s_[self.nghosts : self.n - self.nghosts]

    @property
    def f_(self) -> slice:
        return np.s_[self.nghosts : self.f.numel() - self.nghosts]

    @property
    def b_(self) -> tuple[int, int, int]:
        return (-1, self.n - self.nghosts - 1, self.nghosts)

    @property
    def g_(self) -> tuple[None, slice, slice]:
        return (None, np.s_[self.n - self.nghosts :], np.s_[: self.nghosts])

    @property
    def gi_(self) -> tuple[None, slice, slice]:
        g = self.nghosts
        return (None, np.s_[self.n - 2 * g : self.n - g], np.s_[g : 2 * g])


@dataclass(frozen=True)
class UniformGrid(Grid):
    pass


def make_uniform_cell_grid(
    a: float,
    b: float,
    n: int,
    *,
    nghosts: int = 1,
    dtype: Any = None,
    device: torch.device | str | None = None,
) -> UniformGrid:
    """``grid.py:131-179``. The face coordinates are generated in FP64 by the
    same ``linspace`` as the reference and rounded to ``dtype``."""
    if b < a:
        raise ValueError(f"Incorrect interval a > b: {a!r} > {b!r}.")
    if n <= 0:
        raise ValueError(f"Number of cells should be > 0: {n!r} <= 0.")
    assert nghosts >= 0

    if dtype is None:
        dtype = torch.float64
    np_dtype = np.float64 if dtype == torch.float64 else np.float32
    if device is None:
        device = default_device()

    h = (b - a) / n
    f = np.linspace(a - nghosts * h, b + nghosts * h, n + 2 * nghosts + 1,
                    dtype=np_dtype)
    x = (f[1:] + f[:-1]) / 2
    df = np.diff(x)
    dx = np.full_like(x, h)

    def dev(v: np.ndarray) -> torch.Tensor:
        return torch.from_numpy(np.ascontiguousarray(v)).to(device)

    return UniformGrid(
        a=a, b=b, nghosts=nghosts,
        x=dev(x), dx=dev(dx), f=dev(f), df=dev(df),
        dx_min=torch.tensor(h, dtype=dtype, device=device),
        dx_max=torch.tensor(h, dtype=dtype, device=device),
        is_periodic=False,
        h=float(h),
    )


# {{{ quadrature


@dataclass(frozen=True)
class Quadrature:
    """Composite Gauss-Legendre quadrature (``grid.py:344-390``)."""

    order: int
    x: torch.Tensor
    w: torch.Tensor
    dx: torch.Tensor

    def __post_init__(self) -> None:
        if self.x.shape != self.w.shape:
            raise ValueError(
                "'x' and 'w' should have the same shape: "
                f"got {tuple(self.x.shape)} and {tuple(self.w.shape)}.")

    @property
    def nnodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def ncells(self) -> int:
        return int(self.x.shape[1])

    def __call__(self, fn: Callable[[torch.Tensor], torch.Tensor],
                 axis: int | None = None) -> torch.Tensor:
        if axis not in {0, None}:
            raise ValueError(f"Unsupported axis value: {axis!r}.")
        if axis is None:
            return torch.sum(fn(self.x) * self.w)
        return torch.sum(fn(self.x) * self.w, dim=0)


def make_leggauss_quadrature(grid: Grid, order: int) -> Quadrature:
    """``grid.py:393-423``."""
    if order < 1:
        raise ValueError(f"Invalid order: {order!r}.")

    from numpy.polynomial.legendre import leggauss

    xi, wi = leggauss(order)
    xi = torch.tensor(xi, dtype=grid.dtype, device=grid.device).reshape(-1, 1)
    wi = torch.tensor(wi, dtype=grid.dtype, device=grid.device).reshape(-1, 1)

    x = grid.f
    dx = x[1:] - x[:-1]
    dxm = 0.5 * dx.reshape(1, -1)
    xm = 0.5 * (x[1:] + x[:-1]).reshape(1, -1)

    return Quadrature(order=order, x=xm + dxm * xi, w=dxm * wi, dx=dx)


def cell_average(quad: Quadrature, fn: Callable[[torch.Tensor], torch.Tensor]) -> torch.Tensor:
    """``grid.py:426-435``."""
    return quad(fn, axis=0) / quad.dx


# }}}


# {{{ norms


def norm(grid: Grid, u: float | torch.Tensor, *, p: Any = 1,
         weighted: bool = False) -> torch.Tensor:
    """Interior norms (``grid.py:473-506``), computed by the library's
    reduction kernels for ``p`` in ``{1, 2, inf, "tvd"}``."""
    if not isinstance(u, torch.Tensor) or u.ndim == 0:
        return torch.tensor(abs(float(u)), dtype=grid.dtype, device=grid.device)

    from pyshocks_amd import reductions

    if u.shape[:1] == grid.x.shape:
        lo, hi = grid.nghosts, grid.n - grid.nghosts
        w = grid.h if weighted else 1.0
    elif u.shape[:1] == grid.f.shape:
        lo, hi = grid.nghosts, grid.f.numel() - grid.nghosts
        w = grid.h if weighted else 1.0
        if weighted:
            # df is uniform up to the rounding of linspace
            w = float(grid.df[grid.nghosts].item())
    else:
        raise ValueError(f"Array has unexpected shape: {tuple(u.shape[:1])}")

    return reductions.interior_norm(u, lo, hi, p=p, weight=w)


def rnorm(grid: Grid, u: float | torch.Tensor, v: float | torch.Tensor, *,
          p: Any = 1, weighted: bool = False, atol: float = 1.0e-14) -> torch.Tensor:
    """``grid.py:509-531``."""
    vnorm = norm(grid, v, p=p, weighted=weighted)
    if float(vnorm) < atol:
        vnorm = torch.tensor(1.0, dtype=grid.dtype, device=grid.device)

    return norm(grid, u - v, p=p, weighted=weighted) / vnorm


# }}}
