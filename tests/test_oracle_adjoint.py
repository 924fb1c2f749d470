"""Pins of the complex-step Jacobians (oracle/adjoint_cs.py) used to check
the reverse kernels: on real input they reproduce the oracle's operator and
step, and their columns agree with central finite differences."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import adjoint_cs as A
from oracle import weno_np as W


def _data(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = (np.arange(n + 6) - 2.5) / n
    return 0.3 + np.sin(2 * np.pi * x) + 0.2 * rng.standard_normal(n + 6)


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo", "advection"])
def test_complex_restatement_matches_oracle(kind: str) -> None:
    n = 48
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u = _data(n, 1)
    vel = np.ones(grid.n)
    ref = W.apply_operator(W.Scheme(kind, eps=1e-6, velocity=vel), grid, 0.0, u)
    got = A.apply_operator_cs(kind, grid, u.astype(np.complex128), eps=1e-6)
    np.testing.assert_allclose(got.real[3:-3], ref[3:-3], rtol=0, atol=1e-13)


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo", "advection"])
def test_step_jacobian_vs_central_fd(kind: str) -> None:
    n = 24
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u = _data(n, 2)
    dt = 0.4 / n
    J = A.step_jacobian(kind, grid, u, dt, eps=1e-6)
    vel = np.ones(grid.n)
    sch = W.Scheme(kind, eps=1e-6, velocity=vel)

    def step(x):
        return W.ssprk33_advance(lambda t, v: W.apply_operator(sch, grid, t, v), dt, 0.0,
                                 W.apply_boundary_periodic(grid, x))

    h = 1e-6
    for j in (0, 5, 11, 23):
        e = np.zeros_like(u)
        e[3 + j] = h
        fd = (step(u + e) - step(u - e))[3:-3] / (2 * h)
        np.testing.assert_allclose(J[:, j], fd, rtol=0, atol=1e-7 * max(1.0, np.abs(fd).max()))


def test_constant_state_jacobian_is_linear_scheme() -> None:
    """At a constant state the WENO weights are the ideal ones, so the
    advection operator Jacobian is the linear 5th-order upwind-biased
    difference: rows are shifts of [-2, 15, -60, 20, 30, -3] / (60 dx)."""
    n = 16
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u = np.full(grid.n, 0.7)
    J = A.operator_jacobian("advection", grid, u, eps=1e-6)
    # L_i = -(F_{i+1/2} - F_{i-1/2}) / dx with F_{i+1/2} = [2,-13,47,27,-3]/60
    # on cells i-2..i+2
    c = np.array([2.0, -13.0, 47.0, 27.0, -3.0]) / 60.0
    row = np.zeros(6)
    row[1:6] -= c  # F_{i+1/2} on i-2..i+2 -> offsets -2..2 in slots 1..5
    row[0:5] += c  # F_{i-1/2} on i-3..i+1
    row = row / grid.h if hasattr(grid, "h") else row * n
    for i in range(n):
        expected = np.zeros(n)
        for k, off in enumerate(range(-3, 3)):
            expected[(i + off) % n] += row[k]
        np.testing.assert_allclose(J[i], expected, rtol=0, atol=1e-9 * n)
