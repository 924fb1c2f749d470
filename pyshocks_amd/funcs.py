"""Initial conditions of the workloads: periodic sums of sinusoids plus a
step, with exact cell averages computed on the device.

``u0(x) = c + sum_m A_m sin(2 pi k_m x + phi_m) + H 1[frac(x) >= xs]`` on
``[0, 1)``; its antiderivative is closed-form, so cell averages are exact
differences of the primitive at the faces (the reference computes
Gauss-Legendre averages instead, ``grid.py:393-435``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from pyshocks_amd.grid import Grid


@dataclass(frozen=True)
class SineStep:
    c: float = 0.0
    amplitudes: tuple[float, ...] = ()
    wavenumbers: tuple[int, ...] = ()
    phases: tuple[float, ...] = ()
    step_height: float = 0.0
    step_location: float = 0.5

    def primitive(self, y: torch.Tensor) -> torch.Tensor:
        """``int_0^y u0`` in FP64."""
        y = y.to(torch.float64)
        w = self.c * y
        for a, k, p in zip(self.amplitudes, self.wavenumbers, self.phases):
            w = w + a * (math.cos(p) - torch.cos(2.0 * math.pi * k * y + p)) / (2.0 * math.pi * k)
        if self.step_height != 0.0:
            fl = torch.floor(y)
            frac = y - fl
            w = w + self.step_height * (fl * (1.0 - self.step_location)
                                        + torch.clamp(frac - self.step_location, min=0.0))
        return w

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        x = x.to(torch.float64)
        u = torch.full_like(x, self.c)
        for a, k, p in zip(self.amplitudes, self.wavenumbers, self.phases):
            u = u + a * torch.sin(2.0 * math.pi * k * x + p)
        if self.step_height != 0.0:
            u = u + self.step_height * ((x - torch.floor(x)) >= self.step_location).to(x.dtype)
        return u


def sine_offset(c: float = 0.5) -> SineStep:
    """``u0(x) = c + sin(2 pi x)``."""
    return SineStep(c=c, amplitudes=(1.0,), wavenumbers=(1,), phases=(0.0,))


def random_sine_step(seed: int = 0) -> SineStep:
    """8 sinusoids with amplitudes in [-1, 1], wavenumbers 1..32, random
    phases, plus a step of height in [-1, 1] at a random location (the
    large single-instance workload)."""
    rng = np.random.default_rng(seed)
    amps = rng.uniform(-1.0, 1.0, size=8)
    ks = rng.integers(1, 33, size=8)
    phis = rng.uniform(0.0, 2.0 * np.pi, size=8)
    height = rng.uniform(-1.0, 1.0)
    xs = rng.uniform(0.05, 0.95)
    return SineStep(c=0.0, amplitudes=tuple(float(a) for a in amps),
                    wavenumbers=tuple(int(k) for k in ks),
                    phases=tuple(float(p) for p in phis),
                    step_height=float(height), step_location=float(xs))


def cell_averages(ic: SineStep, grid: Grid) -> torch.Tensor:
    """Exact cell averages on every cell of ``grid`` (ghosts included) in the
    grid's dtype. The grid must be uniform on ``[0, 1]``."""
    n = grid.ncells
    g = grid.nghosts
    h = 1.0 / n
    f = h * torch.arange(-g, n + g + 1, dtype=torch.float64, device=grid.device)
    w = ic.primitive(f)
    return ((w[1:] - w[:-1]) / h).to(grid.dtype).contiguous()
