"""Discrete adjoint of the fused SSP-RK3 step (reverse stage kernels)."""

from __future__ import annotations

from typing import Any

import torch


def ssprk33_step_vjp(source: Any, u: torch.Tensor, p: torch.Tensor, dt: Any) -> torch.Tensor:
    """``J(u)^T p`` for one SSP-RK3 step of the bound operator."""
    raise NotImplementedError("the reverse stage kernels are not built yet")
