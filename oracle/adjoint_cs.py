"""Complex-step derivatives of the oracle's SSP-RK3 step.

TEST INFRASTRUCTURE ONLY (imported by ``tests/``). The step of
``oracle/weno_np.py`` is re-evaluated on complex arrays with the two
non-analytic operations made analytic by branching on the real part
(``|z| -> z sign(Re z)``, ``max(|a|, |b|)`` picks the operand with the larger
real modulus, the upwind switch compares real parts). Then
``Im f(u + i h e_j) / h`` with ``h = 1e-30`` is the exact derivative
column up to rounding, which gives a dense Jacobian of one step for small
grids -- the quantity the reference's ``adjoint_step`` builds with
``jax.jacfwd`` (``timestepping.py:174``) -- without finite-difference error.
Ties of ``max`` and zeros of the switch are measure-zero and avoided by the
tests' random data.
"""

from __future__ import annotations

import numpy as np

from oracle import weno_np as W


def _abs_cs(z: np.ndarray) -> np.ndarray:
    return z * np.sign(z.real)


def _reconstruct_cs(f: np.ndarray, eps: float) -> tuple[np.ndarray, np.ndarray]:
    # the reference's reconstruction is analytic (convolutions, powers,
    # divisions), so the oracle's own functions apply to complex input
    return W.reconstruct_wenojs(f, eps)


def apply_operator_cs(kind: str, grid: W.Grid, u: np.ndarray, *, eps: float,
                      velocity: float = 1.0) -> np.ndarray:
    """``apply_operator`` of ``oracle.weno_np`` on complex input."""
    u = W.apply_boundary_periodic(grid, u)
    ul, ur = _reconstruct_cs(u, eps)
    if kind == "rusanov":
        fl = ul**2 / 2
        fr = ur**2 / 2
        a = _abs_cs(u)
        a = np.where(a[1:].real >= a[:-1].real, a[1:], a[:-1])
        fnum = 0.5 * (fl[1:] + fr[:-1]) - 0.5 * a * (ul[1:] - ur[:-1])
    elif kind == "godunov":
        fl = ul**2 / 2
        fr = ur**2 / 2
        aavg = (ur[:-1] + ul[1:]) / 2
        fnum = np.where(aavg.real > 0, fr[:-1], fl[1:])
    elif kind == "eo":
        # f(max(ur, 0)) + f(min(ul, 0)), the switches on the real parts
        p = np.where(ur[:-1].real > 0, ur[:-1], 0)
        q = np.where(ul[1:].real < 0, ul[1:], 0)
        fnum = p**2 / 2 + q**2 / 2
    elif kind == "advection":
        fnum = ur[:-1] if velocity > 0 else ul[1:]
    else:
        raise NotImplementedError(kind)
    F = np.pad(fnum, 1)
    L = -(F[1:] - F[:-1]) / grid.dx
    if kind == "advection":
        L = velocity * L
    return L


def step_cs(kind: str, grid: W.Grid, u: np.ndarray, dt: float, *, eps: float,
            velocity: float = 1.0) -> np.ndarray:
    """One SSP-RK3 step (``timestepping.py:312-320``) on complex input."""

    def fn(t, x):
        return apply_operator_cs(kind, grid, x, eps=eps, velocity=velocity)

    return W.ssprk33_advance(fn, dt, 0.0, u)


def step_jacobian(kind: str, grid: W.Grid, u: np.ndarray, dt: float, *,
                  eps: float, velocity: float = 1.0, h: float = 1.0e-30) -> np.ndarray:
    """Dense Jacobian of one step restricted to the interior cells:
    ``J[i, j] = d u+_i / d u_j`` for interior ``i, j`` (the periodic ghost
    fill maps every ghost input onto an interior cell)."""
    g = grid.nghosts
    n = grid.n - 2 * g
    J = np.empty((n, n))
    base = W.apply_boundary_periodic(grid, np.asarray(u, dtype=np.float64))
    for j in range(n):
        z = base.astype(np.complex128)
        z[g + j] += 1j * h
        z = W.apply_boundary_periodic(grid, z)
        out = step_cs(kind, grid, z, dt, eps=eps, velocity=velocity)
        J[:, j] = out[g:g + n].imag / h
    return J


def operator_jacobian(kind: str, grid: W.Grid, u: np.ndarray, *, eps: float,
                      velocity: float = 1.0, h: float = 1.0e-30) -> np.ndarray:
    """Dense interior Jacobian of ``apply_operator`` by complex step."""
    g = grid.nghosts
    n = grid.n - 2 * g
    J = np.empty((n, n))
    base = W.apply_boundary_periodic(grid, np.asarray(u, dtype=np.float64))
    for j in range(n):
        z = base.astype(np.complex128)
        z[g + j] += 1j * h
        z = W.apply_boundary_periodic(grid, z)
        out = apply_operator_cs(kind, grid, z, eps=eps, velocity=velocity)
        J[:, j] = out[g:g + n].imag / h
    return J
