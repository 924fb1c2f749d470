"""WENO-JS5 coefficients, smoothness indicators, interpolants and weights.

Mirrors ``weno.py`` of pyshocks: ``Stencil`` (``:55-98``),
``weno_smoothness`` (``:114-140``), ``weno_interp`` (``:143-157``),
``weno_js_32_coefficients`` (``:166-203``), ``weno_js_53_coefficients``
(``:206-244``), ``weno_js_weights`` (``:247-256``) and
``es_weno_parameters`` (``:265-282``). The field-level functions take the
fifth-order stencil; the third-order reconstructions run whole
(``reconstruction.reconstruct``). The coefficient tables are kept as host tensors (they are
the same numbers as the reference's, taps ordered ``(i+2, ..., i-2)``); the
three array functions evaluate them with the library's kernels, which
implement the fifth-order stencil with zero padding outside the array exactly
as the reference's 'same' convolution does.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import torch

from pyshocks_amd import _lib


@dataclass(frozen=True)
class Stencil:
    """Coefficients ``a, b, c, d`` of a WENO reconstruction."""

    a: torch.Tensor | None
    b: torch.Tensor | None
    c: torch.Tensor
    d: torch.Tensor


def weno_js_53_coefficients(dtype: Any = None) -> Stencil:
    """Coefficients of the fifth-order WENO-JS scheme."""
    if dtype is None:
        dtype = torch.float64

    a = torch.tensor([13.0 / 12.0, 1.0 / 4.0], dtype=dtype)
    b = torch.tensor(
        [
            # i + 2, i + 1, i, i - 1, i - 2
            [[0.0, 0.0, 1.0, -2.0, 1.0], [0.0, 0.0, 3.0, -4.0, 1.0]],
            [[0.0, 1.0, -2.0, 1.0, 0.0], [0.0, 1.0, 0.0, -1.0, 0.0]],
            [[1.0, -2.0, 1.0, 0.0, 0.0], [1.0, -4.0, 3.0, 0.0, 0.0]],
        ],
        dtype=dtype,
    )
    c = torch.tensor(
        [
            [0.0, 0.0, 11.0 / 6.0, -7.0 / 6.0, 2.0 / 6.0],
            [0.0, 2.0 / 6.0, 5.0 / 6.0, -1.0 / 6.0, 0.0],
            [-1.0 / 6.0, 5.0 / 6.0, 2.0 / 6.0, 0.0, 0.0],
        ],
        dtype=dtype,
    )
    d = torch.tensor([[1.0 / 10.0, 6.0 / 10.0, 3.0 / 10.0]], dtype=dtype).T

    return Stencil(a=a, b=b, c=c, d=d)


def weno_js_32_coefficients(dtype: Any = None) -> Stencil:
    """Coefficients of the third-order WENO-JS scheme (``weno.py:166-203``),
    taps ordered ``(i+1, i, i-1)``."""
    if dtype is None:
        dtype = torch.float64

    a = torch.tensor([1.0], dtype=dtype)
    b = torch.tensor([[[0.0, 1.0, -1.0]], [[1.0, -1.0, 0.0]]], dtype=dtype)
    c = torch.tensor([[0.0, 3.0 / 2.0, -1.0 / 2.0], [1.0 / 2.0, 1.0 / 2.0, 0.0]],
                     dtype=dtype)
    d = torch.tensor([[1.0 / 3.0, 2.0 / 3.0]], dtype=dtype).T

    return Stencil(a=a, b=b, c=c, d=d)


def es_weno_parameters(grid: Any, u0: torch.Tensor) -> tuple[torch.Tensor, float]:
    """``(eps, delta)`` of ES-WENO (``weno.py:265-282``): ``eps = sum_int
    dx |u0| * dx_min^2``, ``delta = dx_min^2``."""
    from pyshocks_amd.grid import norm

    eps = norm(grid, u0, p=1, weighted=True) * grid.dx_min**2
    return eps, grid.dx_min**2


_JS53 = weno_js_53_coefficients()


def _check_js53(s: Stencil) -> None:
    if s.a is None or s.b is None or s.c.shape != (3, 5):
        raise NotImplementedError(
            "only the fifth-order WENO-JS stencil (weno_js_53_coefficients) is "
            "implemented on the GPU")
    if not (torch.equal(s.a.double(), _JS53.a) and torch.equal(s.b.double(), _JS53.b)
            and torch.equal(s.c.double(), _JS53.c)
            and torch.equal(s.d.double(), _JS53.d)):
        raise NotImplementedError("custom WENO coefficients are not supported")


def weno_smoothness(s: Stencil, u: torch.Tensor, *, mode: Any = None) -> torch.Tensor:
    """``beta`` of shape ``(3, u.numel())``."""
    if mode is not None:
        raise NotImplementedError("only the 'same' convolution mode is supported")
    _check_js53(s)
    _lib.require_device(u)
    beta = torch.empty((3, u.numel()), dtype=u.dtype, device=u.device)
    _lib.call("psa_weno_smoothness", _lib.dtype_code(u.dtype), _lib.ptr(u),
              _lib.ptr(beta), u.numel(), _lib.stream_of(u))
    return beta


def weno_interp(s: Stencil, u: torch.Tensor, *, mode: Any = None) -> torch.Tensor:
    """``uhat`` of shape ``(3, u.numel())``."""
    if mode is not None:
        raise NotImplementedError("only the 'same' convolution mode is supported")
    _check_js53(s)
    _lib.require_device(u)
    uhat = torch.empty((3, u.numel()), dtype=u.dtype, device=u.device)
    _lib.call("psa_weno_interp", _lib.dtype_code(u.dtype), _lib.ptr(u),
              _lib.ptr(uhat), u.numel(), _lib.stream_of(u))
    return uhat


def weno_js_weights(s: Stencil, u: torch.Tensor, *, eps: float) -> torch.Tensor:
    """``omega`` of shape ``(3, u.numel())`` with ``p = 2``."""
    _check_js53(s)
    _lib.require_device(u)
    omega = torch.empty((3, u.numel()), dtype=u.dtype, device=u.device)
    _lib.call("psa_weno_js_weights", _lib.dtype_code(u.dtype), _lib.ptr(u),
              _lib.ptr(omega), u.numel(), float(eps), _lib.stream_of(u))
    return omega
