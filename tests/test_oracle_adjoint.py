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


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo", "advection", "lf"])
def test_complex_restatement_matches_oracle(kind: str) -> None:
    n = 48
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u = _data(n, 1)
    vel = np.ones(grid.n)
    ref = W.apply_operator(W.Scheme(kind, eps=1e-6, velocity=vel), grid, 0.0, u)
    got = A.apply_operator_cs(kind, grid, u.astype(np.complex128), eps=1e-6)
    np.testing.assert_allclose(got.real[3:-3], ref[3:-3], rtol=0, atol=1e-13)


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo", "advection", "lf"])
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


def _learned_data(n: int, seed: int):
    from oracle import mlp_np as M

    rng = np.random.default_rng(seed)
    p = (0.3 * rng.standard_normal(M.NPARAM)).astype(np.float32)
    # W1, W2, W3 representable in BF16, so that the unrounded oracle uses the
    # surrogate's weights
    P = M.unpack(p)
    p = M.pack(M.bf16(P["W1"]), P["b1"], M.bf16(P["W2"]), P["b2"], M.bf16(P["W3"]), P["b3"])
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    x = (np.arange(n + 6) - 2.5) / n
    u = 0.5 + np.sin(2 * np.pi * x) + 0.3 * (x > 0.5) + 0.02 * rng.standard_normal(n + 6)
    return p, grid, W.apply_boundary_periodic(grid, u)


def test_learned_surrogate_values() -> None:
    """The complex surrogate's real part is the oracle's unrounded learned
    reconstruction (equal weights; b3 + ln d in FP32 against FP64 moves the
    weights by ~1e-7); with tau = 0 every varying cell is WENO-JS."""
    from oracle import mlp_np as M

    p, grid, u = _learned_data(48, 3)
    ul, ur = A.learned_reconstruct_cs(p, 20.0, 1.5, u.astype(complex), 1e-6)
    rl, rr = M.learned_reconstruct(p, 20.0, u, 1e-6, sigma=1.5, emulate=False)
    reg = M.regimes(M.stencils(u), 1e-6, 20.0, 1.5)
    assert np.all(np.bincount(reg, minlength=3) > 0)
    np.testing.assert_allclose(ur.real, rr, rtol=0, atol=1e-6)
    np.testing.assert_allclose(ul.real, rl, rtol=0, atol=1e-6)
    jl, jr = W.reconstruct_wenojs(u, 1e-6)
    ul0, ur0 = A.learned_reconstruct_cs(p, 0.0, 0.0, u.astype(complex), 1e-6)
    np.testing.assert_allclose(ur0.real, jr, rtol=0, atol=1e-15)
    np.testing.assert_allclose(ul0.real, jl, rtol=0, atol=1e-15)


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo"])
def test_learned_step_jacobian_vs_fd(kind: str) -> None:
    """Complex-step columns of the learned surrogate's step Jacobian against
    central differences of its real evaluation (the branches fixed by the
    real parts stay fixed for these random data and h = 1e-6)."""
    n = 24
    p, grid, u = _learned_data(n, 4)
    dt = 0.2 / n
    J = A.learned_step_jacobian(p, 20.0, 1.5, kind, grid, u, dt, eps=1e-6)
    h = 1.0e-6
    g = grid.nghosts
    for j in (0, 5, 11, 17, 23):
        e = np.zeros(n + 6)
        e[g + j] = h
        # the real evaluation of the surrogate: a complex step of size 0
        up = _step_real(p, kind, grid, u + e, dt)
        um = _step_real(p, kind, grid, u - e, dt)
        fd = (up - um)[g:g + n] / (2 * h)
        np.testing.assert_allclose(J[:, j], fd, rtol=1e-5, atol=1e-7 * np.abs(J).max())


def _step_real(p, kind, grid, u, dt):
    """One step of the learned surrogate on real input (through the complex
    code with zero imaginary part)."""
    ju = W.apply_boundary_periodic(grid, np.asarray(u, dtype=np.float64))

    def operator(t, x):
        x = W.apply_boundary_periodic(grid, x)
        ul, ur = A.learned_reconstruct_cs(p, 20.0, 1.5, x, 1e-6)
        fl, fr = ul**2 / 2, ur**2 / 2
        if kind == "rusanov":
            a = np.abs(x)
            a = np.maximum(a[1:], a[:-1])
            fnum = 0.5 * (fl[1:] + fr[:-1]) - 0.5 * a * (ul[1:] - ur[:-1])
        else:
            aavg = (ur[:-1] + ul[1:]) / 2
            fnum = np.where(aavg > 0, fr[:-1], fl[1:])
        F = np.pad(fnum, 1)
        return -(F[1:] - F[:-1]) / grid.dx

    out = W.ssprk33_advance(operator, dt, 0.0, ju.astype(np.complex128))
    return out.real


# {{{ the C reverse mode (oracle/weno_ref.c), the CPU adjoint


@pytest.fixture(scope="module")
def ref_c():
    from oracle import build as oracle_build
    from oracle import ref_c as R

    if not R.LIB.exists():
        oracle_build.build(verbose=False)
    return R


def test_c_operator_vjp_is_jacobian_transpose(ref_c) -> None:
    """J_L(u)^T mu of the C reverse mode against the transpose of the
    complex-step operator Jacobian (Rusanov, WENO-JS5), on noisy data that
    exercises every weight and both branches of the Rusanov speed."""
    n = 40
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u = W.apply_boundary_periodic(grid, _data(n, 3))
    mu = np.zeros(grid.n)
    mu[3:-3] = np.random.default_rng(4).standard_normal(n)
    J = A.operator_jacobian("rusanov", grid, u, eps=1e-6)
    got = ref_c.operator_vjp(u, mu, g=3, dx=grid.dx_min, eps=1e-6)
    ref = J.T @ mu[3:-3]
    np.testing.assert_allclose(got[3:-3], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert np.all(got[:3] == 0) and np.all(got[-3:] == 0)


def test_c_step_vjp_is_jacobian_transpose(ref_c) -> None:
    n = 32
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u = W.apply_boundary_periodic(grid, _data(n, 5))
    dt = 0.4 / n
    lam = np.zeros(grid.n)
    lam[3:-3] = np.random.default_rng(6).standard_normal(n)
    J = A.step_jacobian("rusanov", grid, u, dt, eps=1e-6)
    got = ref_c.step_vjp(u, lam, g=3, dx=grid.dx_min, eps=1e-6, dt=dt)
    ref = J.T @ lam[3:-3]
    np.testing.assert_allclose(got[3:-3], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_c_adjoint_is_chained_jacobians(ref_c) -> None:
    """The checkpointed C adjoint (windows of 5 steps over 13, the last one
    partial) against the product of the complex-step Jacobians of every step
    along the oracle's trajectory; J and u(T) against the oracle's solve."""
    n, nsteps, every = 24, 13, 5
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u0 = W.apply_boundary_periodic(grid, _data(n, 7))
    target = W.apply_boundary_periodic(grid, _data(n, 8))
    dt = 0.3 / n
    sch = W.Scheme("rusanov", eps=1e-6)
    states = [u0]
    for _ in range(nsteps):
        x = W.apply_boundary_periodic(grid, states[-1])
        states.append(W.ssprk33_advance(lambda t, v: W.apply_operator(sch, grid, t, v),
                                        dt, 0.0, x))
    uT = states[-1]
    lam = (uT[3:-3] - target[3:-3]) * grid.dx_min
    for k in range(nsteps - 1, -1, -1):
        lam = A.step_jacobian("rusanov", grid, states[k], dt, eps=1e-6).T @ lam
    grad, uf, Jv = ref_c.adjoint_rusanov(u0, target, n=n, g=3, dx=grid.dx_min, eps=1e-6, dt=dt,
                                         nsteps=nsteps, every=every)
    np.testing.assert_allclose(uf[3:-3], uT[3:-3], rtol=0, atol=1e-13)
    np.testing.assert_allclose(Jv, 0.5 * np.sum((uT[3:-3] - target[3:-3]) ** 2) * grid.dx_min,
                               rtol=1e-12)
    np.testing.assert_allclose(grad[3:-3], lam, rtol=0, atol=1e-12 * np.abs(lam).max())
    assert np.all(grad[:3] == 0) and np.all(grad[-3:] == 0)


def test_c_adjoint_dot_product_vs_central_fd(ref_c) -> None:
    """<dJ/du0, d> against a central difference of J along a smooth direction
    at a larger size, through the shock of u0 = 0.5 + sin(2 pi x)."""
    n, nsteps = 256, 120
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    x = grid.x
    u0 = W.apply_boundary_periodic(grid, 0.5 + np.sin(2 * np.pi * x))
    target = W.apply_boundary_periodic(grid, 0.5 + np.sin(2 * np.pi * x + 0.1))
    dt = 0.5 * grid.dx_min / 1.5
    d = np.zeros(grid.n)
    d[3:-3] = np.cos(4 * np.pi * x[3:-3]) + 0.5 * np.sin(6 * np.pi * x[3:-3])
    grad, _, _ = ref_c.adjoint_rusanov(u0, target, n=n, g=3, dx=grid.dx_min, eps=1e-6, dt=dt,
                                       nsteps=nsteps, every=50)

    def J(v: np.ndarray) -> float:
        return ref_c.adjoint_rusanov(v, target, n=n, g=3, dx=grid.dx_min, eps=1e-6, dt=dt,
                                     nsteps=nsteps, every=50)[2]

    h = 1e-6
    fd = (J(u0 + h * d) - J(u0 - h * d)) / (2 * h)
    np.testing.assert_allclose(float(grad @ d), fd, rtol=1e-6)


# }}}
