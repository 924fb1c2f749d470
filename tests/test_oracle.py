"""Pins of the CPU oracle (``oracle/weno_np.py``) that do not depend on it.

Each test checks the restatement against something a plausible mistake
(a dropped term, a flipped tap order, a sign, a wrong index) would violate:
closed-form identities of the coefficients, explicit loops written as in the
reference's own tests, exactness on polynomials, convergence orders, exact
solutions (a shift for advection, Hopf-Lax for Burgers) and invariants.
"""

from __future__ import annotations

from itertools import product

import numpy as np
import pytest

from oracle import exact as E
from oracle import weno_np as W

S = W.WENOJS53


# {{{ coefficients


def test_coefficient_identities() -> None:
    # candidate interpolants reproduce constants
    np.testing.assert_allclose(S.c.sum(axis=1), 1.0, rtol=0, atol=1e-15)
    # ideal weights are a partition of unity
    assert abs(S.d.sum() - 1.0) < 1e-15
    # sum_k d_k c_k is the 5th-order upwind-biased interpolant at x_{i+1/2},
    # taps (i-2 .. i+2) = [2, -13, 47, 27, -3] / 60, i.e. reversed in the
    # (i+2 .. i-2) ordering of the table
    combined = (S.d * S.c).sum(axis=0)
    expected = np.array([-3.0, 27.0, 47.0, -13.0, 2.0]) / 60.0
    np.testing.assert_allclose(combined, expected, rtol=0, atol=1e-15)

    # smoothness indicators vanish on constants and equal s^2 on u_j = s j
    u = np.full(16, 3.25)
    beta = W.weno_smoothness(S, u)
    np.testing.assert_allclose(beta[:, 2:-2], 0.0, atol=1e-14)
    s = 0.7
    u = s * np.arange(16, dtype=np.float64)
    beta = W.weno_smoothness(S, u)
    np.testing.assert_allclose(beta[:, 2:-2], s**2, rtol=1e-13)


def test_flipped_taps_fail_identity() -> None:
    """A transposed tap order must not satisfy the identity above."""
    flipped = (S.d * S.c[:, ::-1]).sum(axis=0)
    expected = np.array([-3.0, 27.0, 47.0, -13.0, 2.0]) / 60.0
    assert np.max(np.abs(flipped - expected)) > 0.1


# }}}


# {{{ loops vs vectorised (tests/test_weno.py:34-108)


def test_smoothness_and_interp_loops() -> None:
    rtol = 2.0e-15
    n = 64
    a, b, c = S.a, S.b, S.c
    nghosts = b.shape[-1] // 2
    stencil = np.arange(-nghosts, nghosts + 1)
    n = n + 2 * nghosts
    m = np.s_[nghosts:-nghosts]

    u = np.sin(np.linspace(0.0, 2.0 * np.pi, n))

    beta0 = np.zeros((3, n))
    for j in range(*m.indices(n)):
        for i, k in product(range(3), range(a.size)):
            beta0[i, j] += a[k] * np.sum(u[j + stencil] * b[i, k, ::-1]) ** 2
    beta1 = W.weno_smoothness(S, u)

    uhat0 = np.zeros((3, n))
    for j in range(*m.indices(n)):
        for i in range(3):
            uhat0[i, j] = np.sum(u[j + stencil] * c[i, ::-1])
    uhat1 = W.weno_interp(S, u)

    err = np.linalg.norm(beta0[:, m] - beta1[:, m]) / np.linalg.norm(beta0[:, m])
    assert err < rtol
    err = np.linalg.norm(uhat0[:, m] - uhat1[:, m]) / np.linalg.norm(uhat0[:, m])
    assert err < rtol


# }}}


# {{{ polynomial exactness


def _poly_cell_averages(coeffs: np.ndarray, f: np.ndarray) -> np.ndarray:
    P = np.polynomial.polynomial.Polynomial(coeffs).integ()
    return (P(f[1:]) - P(f[:-1])) / (f[1:] - f[:-1])


@pytest.mark.parametrize("degree", [0, 1, 2])
def test_candidates_exact_on_quadratics(degree: int) -> None:
    rng = np.random.default_rng(degree)
    coeffs = rng.uniform(-1, 1, size=degree + 1)
    f = np.linspace(0.0, 2.0, 21)
    u = _poly_cell_averages(coeffs, f)
    p = np.polynomial.polynomial.Polynomial(coeffs)

    uhat = W.weno_interp(S, u)
    for k in range(3):
        np.testing.assert_allclose(uhat[k, 2:-2], p(f[1:])[2:-2], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("degree", [3, 4])
def test_combined_exact_on_quartics(degree: int) -> None:
    rng = np.random.default_rng(10 + degree)
    coeffs = rng.uniform(-1, 1, size=degree + 1)
    f = np.linspace(0.0, 2.0, 21)
    u = _poly_cell_averages(coeffs, f)
    p = np.polynomial.polynomial.Polynomial(coeffs)

    uhat = W.weno_interp(S, u)
    right = np.sum(S.d * uhat, axis=0)
    np.testing.assert_allclose(right[2:-2], p(f[1:])[2:-2], rtol=1e-12, atol=1e-13)

    # the left states are the mirrored reconstruction at x_{i-1/2}
    uhat_l = W.weno_interp(S, u[::-1])
    left = np.sum(S.d * uhat_l, axis=0)[::-1]
    np.testing.assert_allclose(left[2:-2], p(f[:-1])[2:-2], rtol=1e-12, atol=1e-13)

    # a single candidate is only third order: it must not be exact here
    assert np.max(np.abs(uhat[0, 2:-2] - p(f[1:])[2:-2])) > 1e-8


# }}}


# {{{ weights (tests/test_weno.py:117-166)


@pytest.mark.parametrize("is_smooth", [True, False])
def test_weights_smooth_and_step(is_smooth: bool) -> None:
    n = 128 + 6
    m = np.s_[3:-3]
    theta = np.linspace(0.0, 2.0 * np.pi, n)
    u = np.sin(theta) if is_smooth else (theta < np.pi).astype(np.float64)
    omega = W.weno_js_weights(S, u, eps=1.0e-12)
    error = np.max(np.abs(omega[:, m] - S.d))
    if is_smooth:
        assert error < 0.1
    else:
        assert error > 0.1


# }}}


# {{{ reconstruction order (tests/test_weno.py:288-338)


def test_reconstruction_order() -> None:
    hs, el, er = [], [], []
    for n in range(32, 256 + 1, 32):
        grid = W.make_uniform_cell_grid(-1.0, 1.0, n, nghosts=3)
        u0 = W.leggauss_cell_average(grid, lambda x: np.sin(2 * np.pi * x), order=6)
        ul, ur = W.reconstruct_wenojs(u0, 1.0e-12)
        fx = np.sin(2 * np.pi * grid.f)
        el.append(W.rnorm(grid, ul, fx[:-1], p=np.inf))
        er.append(W.rnorm(grid, ur, fx[1:], p=np.inf))
        hs.append(grid.dx_max)

    assert W.estimate_order_of_convergence(np.array(hs), np.array(el)) >= 4.5
    assert W.estimate_order_of_convergence(np.array(hs), np.array(er)) >= 4.5


# }}}


# {{{ operator invariants


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo", "advection"])
def test_operator_conservation_and_constants(kind: str) -> None:
    n = 64
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    rng = np.random.default_rng(1)
    vel = np.ones(grid.n)
    sch = W.Scheme(kind, eps=1e-6, velocity=vel)

    u = rng.uniform(-1, 1, size=grid.n)
    L = W.apply_operator(sch, grid, 0.0, u)
    # periodic telescoping: sum_int L dx = 0
    assert abs(np.sum(L[grid.i_]) / n) < 1e-14

    u = np.full(grid.n, 0.75)
    L = W.apply_operator(sch, grid, 0.0, u)
    np.testing.assert_allclose(L[grid.i_], 0.0, atol=1e-12)


def _js_side(v: np.ndarray, eps: float) -> float:
    # v = (u_{i-2}, ..., u_{i+2}), Jiang-Shu (2.17), (2.11)-(2.13), (2.15)
    b0 = 13 / 12 * (v[0] - 2 * v[1] + v[2]) ** 2 + 1 / 4 * (v[0] - 4 * v[1] + 3 * v[2]) ** 2
    b1 = 13 / 12 * (v[1] - 2 * v[2] + v[3]) ** 2 + 1 / 4 * (v[1] - v[3]) ** 2
    b2 = 13 / 12 * (v[2] - 2 * v[3] + v[4]) ** 2 + 1 / 4 * (3 * v[2] - 4 * v[3] + v[4]) ** 2
    q0 = (2 * v[0] - 7 * v[1] + 11 * v[2]) / 6
    q1 = (-v[1] + 5 * v[2] + 2 * v[3]) / 6
    q2 = (2 * v[2] + 5 * v[3] - v[4]) / 6
    a = np.array([0.1 / (eps + b0) ** 2, 0.6 / (eps + b1) ** 2, 0.3 / (eps + b2) ** 2])
    w = a / a.sum()
    return float(w @ np.array([q0, q1, q2]))


def test_rusanov_closed_form_small() -> None:
    """Brute-force flux on a tiny grid with hand-evaluated formulas."""
    n = 8
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    rng = np.random.default_rng(7)
    u = W.apply_boundary_periodic(grid, rng.uniform(-1, 1, size=grid.n))
    eps = 1e-6

    F = W.scalar_flux_rusanov(grid, u, u, eps=eps)
    for j in range(3, grid.n - 2):  # faces with full stencils
        ur = _js_side(u[j - 3: j + 2], eps)
        ul = _js_side(u[j - 2: j + 3][::-1], eps)
        s = max(abs(u[j - 1]), abs(u[j]))
        ref = 0.5 * (ul**2 / 2 + ur**2 / 2) - 0.5 * s * (ul - ur)
        assert abs(F[j] - ref) < 1e-13


def test_engquist_osher_integral_definition() -> None:
    """The Engquist-Osher face flux is (f(a) + f(b)) / 2 - 1/2 int_a^b |f'| du
    (LeVeque (12.49)); for f = u^2 / 2, int_a^b |u| du = (b|b| - a|a|) / 2.
    Checked against that closed form and against midpoint quadrature."""
    rng = np.random.default_rng(11)
    a = rng.uniform(-2, 2, size=200)
    b = rng.uniform(-2, 2, size=200)
    a[:4] = [0.0, 1.0, -1.0, 0.5]
    b[:4] = [0.0, -1.0, 1.0, 0.5]
    F = W.engquist_osher_faces(a, b)
    closed = (a**2 / 2 + b**2 / 2) / 2 - 0.5 * (b * np.abs(b) - a * np.abs(a)) / 2
    np.testing.assert_allclose(F, closed, rtol=0, atol=1e-14)

    k = 20000
    for ai, bi, fi in zip(a[:20], b[:20], F[:20]):
        s = (np.arange(k) + 0.5) / k
        quad = np.sum(np.abs(ai + (bi - ai) * s)) * (bi - ai) / k
        ref = (ai**2 / 2 + bi**2 / 2) / 2 - 0.5 * quad
        assert abs(fi - ref) < 1e-7
    # consistency
    np.testing.assert_allclose(W.engquist_osher_faces(a, a), a**2 / 2, rtol=0,
                               atol=1e-15)


def test_engquist_osher_uses_reconstructed_sides() -> None:
    """The left-of-face state is the right reconstruction of the cell to the
    left, the right-of-face state the left reconstruction of the cell to the
    right (brute force with hand-evaluated WENO-JS formulas)."""
    n = 8
    eps = 1e-6
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u = W.apply_boundary_periodic(
        grid, np.random.default_rng(5).uniform(-1, 1, size=grid.n))
    F = W.scalar_flux_engquist_osher(grid, u, eps=eps)
    for j in range(3, grid.n - 2):
        ur = _js_side(u[j - 3: j + 2], eps)
        ul = _js_side(u[j - 2: j + 3][::-1], eps)
        ref = max(ur, 0.0) ** 2 / 2 + min(ul, 0.0) ** 2 / 2
        assert abs(F[j] - ref) < 1e-13


def test_lax_friedrichs_global_speed() -> None:
    """Lax-Friedrichs is the Rusanov flux with one speed max|a| over the whole
    array as given, ghost entries included."""
    n = 16
    eps = 1e-6
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    rng = np.random.default_rng(3)
    # |u| constant: the local and the global speeds agree
    u = 0.8 * rng.choice([-1.0, 1.0], size=grid.n)
    np.testing.assert_allclose(W.scalar_flux_lax_friedrichs(grid, u, u, eps=eps),
                               W.scalar_flux_rusanov(grid, u, u, eps=eps),
                               rtol=0, atol=1e-15)
    # the largest |u| sits in a ghost entry: every face sees it
    u = rng.uniform(-1, 1, size=grid.n)
    u[0] = -5.0
    F = W.scalar_flux_lax_friedrichs(grid, u, u, eps=eps)
    for j in range(3, grid.n - 2):
        ur = _js_side(u[j - 3: j + 2], eps)
        ul = _js_side(u[j - 2: j + 3][::-1], eps)
        ref = 0.5 * (ul**2 / 2 + ur**2 / 2) - 0.5 * 5.0 * (ul - ur)
        assert abs(F[j] - ref) < 1e-13
    assert F[0] == 0.0 and F[-1] == 0.0
    # alpha != 1: nu = df^(alpha - 1) multiplies the dissipation
    Fa = W.scalar_flux_lax_friedrichs(grid, u, u, eps=eps, alpha=0.9)
    nu = grid.df ** (0.9 - 1)
    d = (F - Fa)[1:-1]
    ul, ur = W.reconstruct_wenojs(u, eps)
    np.testing.assert_allclose(d, -0.5 * 5.0 * (1 - nu) * (ul[1:] - ur[:-1]),
                               rtol=1e-12, atol=1e-14)


# }}}


# {{{ time stepping


def test_ssprk33_order() -> None:
    """``tests/test_timestepping.py:26-99`` for SSPRK33: u' = exp(-t)."""

    def source(t, u):
        return np.array([np.exp(-t)])

    tfinal = 4.0
    u0 = np.array([-1.0])
    u_ex = -np.exp(-tfinal)
    hs, errs = [], []
    for n in range(2, 7):
        maxit, dt = W.predict_maxit_from_timestep(tfinal, 1.0 / 2.0**n)
        event = None
        for event in W.step(lambda t, u, dt=dt: dt, source, u0, maxit=maxit):
            pass
        errs.append(abs(event.u[0] - u_ex) / abs(u_ex))
        hs.append(dt)
    assert W.estimate_order_of_convergence(np.array(hs), np.array(errs)) >= 2.9


def test_step_dt_rule() -> None:
    """Every step adds 1e-15 after clipping to tfinal - t."""
    n = 64
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u0 = E.exact_cell_averages(E.c1_initial_condition(), n)
    sch = W.Scheme("rusanov", eps=1e-6)
    u, dts, t = W.solve(sch, grid, u0, theta=1.0, tfinal=0.1)
    assert t >= 0.1
    assert abs(t - 0.1) < 1e-14
    # all but the last step are CFL steps of 0.5 dx / max|u|, plus 1e-15
    assert all(d > 1e-15 for d in dts)
    assert dts[-1] <= dts[0] + 1e-15


# }}}


# {{{ end-to-end against exact solutions


def test_weno3_coefficients_and_loops() -> None:
    """WENO-JS3 (``weno.py:166-203``): candidates reproduce constants, the
    ideal combination is the third-order interpolant ``-1/6, 5/6, 1/3`` of
    ``(u_{i-1}, u_i, u_{i+1})``, beta is ``s^2`` on ``u_j = s j``, and the
    convolutions equal explicit loops over the taps."""
    S3 = W.WENOJS32
    np.testing.assert_allclose(S3.c.sum(axis=1), 1.0, rtol=0, atol=1e-15)
    combined = (S3.d * S3.c).sum(axis=0)  # taps (i+1, i, i-1)
    np.testing.assert_allclose(combined, [1 / 3, 5 / 6, -1 / 6], rtol=0, atol=1e-15)
    j = np.arange(12, dtype=np.float64)
    beta = W.weno_smoothness(S3, 0.75 * j)
    np.testing.assert_allclose(beta[:, 1:-1], 0.75**2, rtol=1e-14)

    n = 40
    u = np.sin(np.linspace(0.0, 2.0 * np.pi, n))
    beta = W.weno_smoothness(S3, u)
    uhat = W.weno_interp(S3, u)
    for i in range(1, n - 1):
        assert abs(beta[0, i] - (u[i] - u[i - 1]) ** 2) < 1e-15
        assert abs(beta[1, i] - (u[i + 1] - u[i]) ** 2) < 1e-15
        assert abs(uhat[0, i] - (1.5 * u[i] - 0.5 * u[i - 1])) < 2e-15
        assert abs(uhat[1, i] - (0.5 * u[i + 1] + 0.5 * u[i])) < 2e-15


@pytest.mark.parametrize("rec", ["wenojs32", "esweno32"])
def test_weno3_exact_on_linear_data(rec: str) -> None:
    """Each candidate is exact on the cell averages of a linear function, so
    any convex weights (JS or ES) reconstruct it exactly at every face."""
    n = 24
    f = np.linspace(-1.0, 1.0, n + 1)
    u = _poly_cell_averages(np.array([0.3, -0.7]), f)
    ul, ur = W.reconstruct(u, 1.0e-6, rec)
    np.testing.assert_allclose(ur[1:-1], 0.3 - 0.7 * f[2:-1], rtol=0, atol=1e-13)
    np.testing.assert_allclose(ul[1:-1], 0.3 - 0.7 * f[1:-2], rtol=0, atol=1e-13)


def test_esweno_tau_padding() -> None:
    """``weno.py:293``: tau is zero for the first and last entry, whatever the
    data; with tau = 0 the ES weights are the ideal ones."""
    rng = np.random.default_rng(2)
    u = rng.uniform(-1, 1, size=9)
    om = W.es_weno_weights(W.WENOJS32, u, eps=1e-6)
    np.testing.assert_allclose(om[:, 0], [1 / 3, 2 / 3], rtol=1e-15)
    np.testing.assert_allclose(om[:, -1], [1 / 3, 2 / 3], rtol=1e-15)
    tau = (u[2] - 2 * u[1] + u[0]) ** 2
    b0, b1 = (u[1] - u[0]) ** 2, (u[2] - u[1]) ** 2
    a0, a1 = (1 + tau / (1e-6 + b0)) / 3, 2 * (1 + tau / (1e-6 + b1)) / 3
    np.testing.assert_allclose(om[:, 1], [a0 / (a0 + a1), a1 / (a0 + a1)], rtol=1e-14)


@pytest.mark.parametrize("rec", ["wenojs32", "esweno32"])
def test_weno3_reconstruction_order(rec: str) -> None:
    """``tests/test_weno.py:288-338`` for the third-order reconstructions:
    order >= 2.5 on a smooth sine (N = 192 .. 384)."""
    hs, el, er = [], [], []
    for n in range(192, 384 + 1, 32):
        grid = W.make_uniform_cell_grid(-1.0, 1.0, n, nghosts=2)
        u0 = W.leggauss_cell_average(grid, lambda x: np.sin(2 * np.pi * x), order=4)
        ul, ur = W.reconstruct(u0, 1.0e-6, rec)
        fx = np.sin(2 * np.pi * grid.f)
        el.append(W.rnorm(grid, ul, fx[:-1], p=np.inf))
        er.append(W.rnorm(grid, ur, fx[1:], p=np.inf))
        hs.append(grid.dx_max)

    assert W.estimate_order_of_convergence(np.array(hs), np.array(el)) >= 2.5
    assert W.estimate_order_of_convergence(np.array(hs), np.array(er)) >= 2.5


def test_weno3_brute_force_faces() -> None:
    """Both face states of every cell on a tiny grid from the Jiang-Shu
    third-order formulas, for both weight families."""
    eps = 1e-6
    rng = np.random.default_rng(8)
    u = rng.uniform(-1, 1, size=10)
    for rec in ("wenojs32", "esweno32"):
        ul, ur = W.reconstruct(u, eps, rec)
        for i in range(1, u.size - 1):
            um, u0, up = u[i - 1], u[i], u[i + 1]
            tau = (up - 2 * u0 + um) ** 2

            def side(a: float, b: float, c: float) -> float:
                # a = u_{i-1}, b = u_i, c = u_{i+1} (mirrored for the left side)
                b0, b1 = (b - a) ** 2, (c - b) ** 2
                if rec == "wenojs32":
                    w0, w1 = (1 / 3) / (eps + b0) ** 2, (2 / 3) / (eps + b1) ** 2
                else:
                    w0, w1 = (1 + tau / (eps + b0)) / 3, 2 * (1 + tau / (eps + b1)) / 3
                return (w0 * (1.5 * b - 0.5 * a) + w1 * (0.5 * b + 0.5 * c)) / (w0 + w1)

            assert abs(ur[i] - side(um, u0, up)) < 1e-14
            assert abs(ul[i] - side(up, u0, um)) < 1e-14


@pytest.mark.parametrize("rec", ["wenojs32", "esweno32"])
def test_advection_order_weno3(rec: str) -> None:
    """``tests/test_convergence.py:400-402`` (godunov + wenojs32 / esweno32,
    order 3 - 0.5), periodic on [-1, 1] to t = 1 with dt = 8 (2/n)^(5/3)."""
    a, b, tfinal = -1.0, 1.0, 1.0

    def ic(grid: W.Grid, x: np.ndarray) -> np.ndarray:
        xm = (grid.b + grid.a) / 2
        xx = (x - xm) / (grid.b - grid.a)
        return np.sin(2 * np.pi * xx + np.sin(2 * np.pi * xx) / np.pi)

    hs, errs = [], []
    for n in range(192, 384 + 1, 64):
        grid = W.make_uniform_cell_grid(a, b, n, nghosts=3)
        vel = np.ones(grid.n)
        sch = W.Scheme("advection", eps=1.0e-6, velocity=vel, rec=rec)
        u0 = W.leggauss_cell_average(grid, lambda x: ic(grid, x), order=4)
        uhat = W.leggauss_cell_average(grid, lambda x: ic(grid, x - tfinal), order=4)
        maxit, dt = W.predict_maxit_from_timestep(tfinal, 8.0 * ((b - a) / n) ** (5.0 / 3.0))
        u, _, _ = W.solve(sch, grid, u0, theta=1.0, maxit=maxit, fixed_dt=dt)
        errs.append(W.norm(grid, u - uhat, p=1, weighted=True)
                    / W.norm(grid, uhat, p=1, weighted=True))
        hs.append(float(np.max(np.diff(grid.f))))
    assert W.estimate_order_of_convergence(np.array(hs), np.array(errs)) >= 2.5


LIMITERS = [("minmod", 1.0), ("minmod", 1.5), ("minmod", 2.0), ("mc", 0.0),
            ("superbee", 0.0), ("vanalbada", 1.0), ("vanalbada", 2.0), ("vanleer", 0.0),
            ("koren", 0.0)]


@pytest.mark.parametrize(("name", "param"), LIMITERS)
def test_limiter_values_and_symmetry(name: str, param: float) -> None:
    """Closed forms of ``limiters.py``: phi(1) = 1 (second order on linear
    data), phi(r) = 0 for r <= 0 (for van Albada 1, (r^2 + r) / (r^2 + 1),
    only on [-1, 0]), Sweby's TVD region 0 <= phi(r) <= min(2, 2 r), and
    phi(1/r) = phi(r) / r for the limiters flagged symmetric (the shortcut
    the MUSCL reconstruction takes)."""
    r = np.linspace(0.01, 5.0, 500)
    phi = W.limiter_phi(name, param, r)
    if not (name == "vanalbada" and param == 2):
        assert abs(float(W.limiter_phi(name, param, np.array(1.0))) - 1.0) < 1e-15
    neg = -r[r <= 1.0] if (name == "vanalbada" and param == 1) else -r
    np.testing.assert_array_equal(W.limiter_phi(name, param, neg), 0.0)
    assert np.all(phi >= 0) and np.all(phi <= np.minimum(2, 2 * r) + 1e-15)
    sym = np.allclose(W.limiter_phi(name, param, 1 / r), phi / r, rtol=1e-13, atol=0)
    assert sym == W.limiter_symmetric(name, param)


@pytest.mark.parametrize(("name", "param"), LIMITERS)
def test_muscl_linear_data_and_steps(name: str, param: float) -> None:
    """MUSCL reproduces the face values of linear data wherever phi(1) = 1,
    drops to first order (the cell value) on a piecewise-constant step
    (``tests/test_muscl.py:60-95``), pads ``ur`` of the last and ``ul`` of
    the first cell with zero, and equals a brute-force loop of the
    reconstruction formulas on random data."""
    n = 16
    f = np.linspace(-1.0, 1.0, n + 1)
    u = 0.25 + 0.5 * (f[1:] + f[:-1]) / 2
    ul, ur = W.reconstruct_muscl(u, name, param)
    if not (name == "vanalbada" and param == 2):
        np.testing.assert_allclose(ur[1:-1], 0.25 + 0.5 * f[2:-1], rtol=0, atol=1e-14)
        np.testing.assert_allclose(ul[1:-1], 0.25 + 0.5 * f[1:-2], rtol=0, atol=1e-14)
    assert ur[-1] == 0.0 and ul[0] == 0.0

    step = np.where(np.arange(n) < n // 2, 1.0, 0.0)
    ul, ur = W.reconstruct_muscl(step, name, param)
    # non-symmetric limiters evaluate phi(atol) where 1/r is cut off, so the
    # first-order value holds to within atol there
    tol = 0.0 if W.limiter_symmetric(name, param) else 1.0e-11
    np.testing.assert_allclose(ur[:-1], step[:-1], rtol=0, atol=tol)
    np.testing.assert_allclose(ul[1:], step[1:], rtol=0, atol=tol)

    u = np.random.default_rng(4).uniform(-1, 1, size=n)
    ul, ur = W.reconstruct_muscl(u, name, param)
    atol = 1.0e-12
    for i in range(1, n - 1):
        sl, sr = u[i] - u[i - 1], u[i + 1] - u[i]
        r = 0.0 if abs(sr) < atol else sl / sr
        phi = float(W.limiter_phi(name, param, np.array(r)))
        if W.limiter_symmetric(name, param):
            phi_inv = 0.0 if abs(r) < atol else phi / r
        else:
            phi_inv = float(W.limiter_phi(name, param,
                                          np.array(atol if abs(r) < atol else 1 / r)))
        assert abs(ur[i] - (u[i] + 0.5 * phi * sr)) < 1e-15
        assert abs(ul[i] - (u[i] - 0.5 * phi_inv * sl)) < 1e-15


SLOPE_LIMITERS = [("minmod", 1.0), ("mc", 0.0), ("superbee", 0.0), ("vanalbada", 1.0),
                  ("vanalbada", 2.0), ("vanleer", 0.0), ("none", 0.0)]


@pytest.mark.parametrize(("name", "param", "du"), [
    ("minmod", 1.0, 1.0), ("mc", 0.0, 1.0), ("superbee", 0.0, 2.0),
    ("vanalbada", 1.0, 1.2), ("vanalbada", 2.0, 0.6), ("vanleer", 0.0, 1.5),
    ("none", 0.0, 3.0)])
def test_slope_limit_known_answers(name: str, param: float, du: float) -> None:
    """Slopes worked by hand for u = (1, 2, 5) on unit cells: left slope 1,
    right slope 3, central 2, and the unlimited slope of ``limiters.py:149``
    is the sum (1 + 5) / 2; the end cells are zero."""
    x = np.array([0.5, 1.5, 2.5])
    s = W.slope_limit(name, param, x, np.array([1.0, 2.0, 5.0]))
    assert s[0] == 0.0 and s[2] == 0.0
    assert abs(s[1] - du) < 1e-15


@pytest.mark.parametrize(("name", "param"), SLOPE_LIMITERS)
def test_muscls_linear_data_and_extrema(name: str, param: float) -> None:
    """Every limited slope of linear data is its slope, so the faces are
    exact (van Albada 2 gives 1, the unlimited form the centred sum);
    at a local extremum every limited slope is zero (first order)."""
    n = 16
    grid = W.make_uniform_cell_grid(-1.0, 1.0, n, nghosts=2)
    x = grid.x
    u = 0.25 + 0.5 * x
    ul, ur = W.reconstruct_muscls(u, name, param, x)
    slope = {"vanalbada": 0.5 if param == 1 else 1.0, "none": None}.get(name, 0.5)
    if slope is None:
        slope = (u[2:] + u[:-2]) / (x[2:] - x[:-2])
    h = grid.dx[grid.i_][0] / 2
    np.testing.assert_allclose(ur[1:-1], u[1:-1] + h * slope, rtol=0, atol=1e-14)
    np.testing.assert_allclose(ul[1:-1], u[1:-1] - h * slope, rtol=0, atol=1e-14)
    assert ul[0] == ur[0] == u[0] and ul[-1] == ur[-1] == u[-1]

    if name != "none":
        peak = np.array([0.0, 1.0, 3.0, 2.0, 1.0])
        ul, ur = W.reconstruct_muscls(peak, name, param, np.arange(5.0))
        assert ul[2] == ur[2] == 3.0


@pytest.mark.parametrize(("name", "param"), [("minmod", 1.0), ("superbee", 0.0),
                                             ("vanalbada", 1.0), ("vanleer", 0.0)])
def test_muscls_matches_flux_limiter_form(name: str, param: float) -> None:
    """Where both slopes are positive the slope form of these limiters is
    phi(r) times the right slope, so the slope-limited faces equal the
    flux-limited MUSCL faces (``reconstruction.py:205-252``); for minmod and
    van Leer that holds in every interior cell. Superbee's slope form takes
    the larger of two signed minmods and van Albada 1's is zero for r < 0,
    so they differ from the flux forms elsewhere; MC's slope form also
    limits by the central slope."""
    n = 200
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=2)
    u = np.random.default_rng(11).uniform(-1, 1, size=grid.x.size)
    ul, ur = W.reconstruct_muscls(u, name, param, grid.x)
    vl, vr = W.reconstruct_muscl(u, name, param)
    d = np.diff(u)
    up = np.zeros(u.size, dtype=bool)
    up[1:-1] = (d[:-1] > 0) & (d[1:] > 0)
    every = name in ("minmod", "vanleer")
    mask = np.ones(u.size, dtype=bool) if every else up
    mask[0] = mask[-1] = False
    assert np.count_nonzero(mask & up) > 20
    np.testing.assert_allclose(ur[mask], vr[mask], rtol=0, atol=1e-13)
    np.testing.assert_allclose(ul[mask], vl[mask], rtol=0, atol=1e-13)
    if not every:
        assert np.max(np.abs(ur - vr)[1:-1]) > 1e-3
    ml, mr = W.reconstruct_muscls(u, "mc", 0.0, grid.x)
    wl, wr = W.reconstruct_muscl(u, "mc", 0.0)
    assert np.max(np.abs(mr - wr)[1:-1]) > 1e-3


@pytest.mark.parametrize(("name", "param"), [("minmod", 1.0), ("vanleer", 0.0),
                                             ("superbee", 0.0)])
def test_muscl_step_is_tvd(name: str, param: float) -> None:
    """Periodic Burgers with MUSCL + Rusanov at Courant 0.5 from a step: the
    total variation never grows (the TVD property that
    ``tests/test_muscl.py:175-296`` checks for the slope-limited form)."""
    n = 128
    grid = W.make_uniform_cell_grid(-1.0, 1.0, n, nghosts=2)
    x = grid.x
    u0 = W.apply_boundary_periodic(grid, np.where(np.abs(x) < 0.5, 1.0, -0.5))
    sch = W.Scheme("rusanov", eps=0.0, rec=f"muscl-{name}-{param}")

    def tv(u):
        v = u[grid.i_]
        return float(np.sum(np.abs(np.diff(np.append(v, v[0])))))

    tvs = []
    for ev in W.step(lambda t, u: W.predict_timestep(sch, grid, t, u),
                     lambda t, u: W.apply_operator(sch, grid, t, u), u0, maxit=60):
        tvs.append(tv(ev.u))
    assert all(b <= a + 1e-12 for a, b in zip(tvs, tvs[1:]))


@pytest.mark.parametrize("side", [-1, +1])
def test_boundary_fills(side: int) -> None:
    """``scalar.py:441-500``: Dirichlet ghost cells hold g(t, x) of their
    centres; a Neumann fill makes the difference quotient of every mirrored
    pair (ghost, interior) equal to -side g(t) (the reference's sign); the
    interior is untouched."""
    n = 12
    grid = W.make_uniform_cell_grid(-1.0, 1.0, n, nghosts=3)
    u = np.random.default_rng(6).uniform(-1, 1, size=grid.n)
    ud = W.apply_boundary_dirichlet(grid, u, 0.3, side, lambda t, x: t + x**2)
    gs = np.s_[:3] if side == -1 else np.s_[-3:]
    np.testing.assert_array_equal(ud[gs], 0.3 + grid.x[gs] ** 2)
    np.testing.assert_array_equal(ud[grid.i_], u[grid.i_])

    un = W.apply_boundary_neumann(grid, u, 0.3, side, lambda t: 2.5 * t)
    np.testing.assert_array_equal(un[grid.i_], u[grid.i_])
    for k in range(3):
        if side == -1:
            i_g, i_in = 2 - k, 3 + k
        else:
            i_g, i_in = grid.n - 1 - k, grid.n - 6 + k
        q = (un[i_in] - un[i_g]) / (grid.x[i_in] - grid.x[i_g])
        assert abs(q - (-side) * 0.75) < 1e-12


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "eo", "lf"])
def test_operator_constant_state_with_boundaries(kind: str) -> None:
    """A constant state with a matching Dirichlet value or a zero Neumann
    gradient is steady: L vanishes in the interior."""
    n = 32
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u = np.full(grid.n, 0.6)
    sch = W.Scheme(kind, eps=1e-6)

    def dirichlet(g, t, v):
        v = W.apply_boundary_dirichlet(g, v, t, -1, lambda t, x: 0.6 + 0 * x)
        return W.apply_boundary_dirichlet(g, v, t, +1, lambda t, x: 0.6 + 0 * x)

    def neumann(g, t, v):
        v = W.apply_boundary_neumann(g, v, t, -1, lambda t: 0.0)
        return W.apply_boundary_neumann(g, v, t, +1, lambda t: 0.0)

    for bc in (dirichlet, neumann):
        L = W.apply_operator(sch, grid, 0.0, u, bc=bc)
        np.testing.assert_allclose(L[grid.i_], 0.0, atol=1e-12)


def test_advection_order_exact_shift() -> None:
    """``tests/test_convergence.py:395-460`` (godunov + wenojs53), periodic on
    [-1, 1] to t = 1 with dt = 8 (2/n)^(5/3)."""
    a, b, tfinal = -1.0, 1.0, 1.0

    def ic(grid: W.Grid, x: np.ndarray) -> np.ndarray:
        xm = (grid.b + grid.a) / 2
        xx = (x - xm) / (grid.b - grid.a)
        return np.sin(2 * np.pi * xx + np.sin(2 * np.pi * xx) / np.pi)

    hs, errs = [], []
    for n in range(32, 256 + 1, 32):
        grid = W.make_uniform_cell_grid(a, b, n, nghosts=3)
        vel = np.ones(grid.n)
        sch = W.Scheme("advection", eps=1.0e-12, velocity=vel)
        u0 = W.leggauss_cell_average(grid, lambda x: ic(grid, x), order=5)
        uhat = W.leggauss_cell_average(grid, lambda x: ic(grid, x - tfinal), order=5)
        maxit, dt = W.predict_maxit_from_timestep(tfinal, 8.0 * ((b - a) / n) ** (5.0 / 3.0))
        u, _, _ = W.solve(sch, grid, u0, theta=1.0, maxit=maxit, fixed_dt=dt)
        errs.append(W.norm(grid, u - uhat, p=1, weighted=True)
                    / W.norm(grid, uhat, p=1, weighted=True))
        hs.append(float(np.max(np.diff(grid.f))))
    assert W.estimate_order_of_convergence(np.array(hs), np.array(errs)) >= 4.5


def test_hopf_lax_exact_is_exact_before_shock() -> None:
    """The Hopf-Lax primitive reproduces the characteristic solution for a
    smooth state (checks the exact-solution helper itself)."""
    ic = E.c1_initial_condition()
    t = 0.1  # before the shock time 1 / (2 pi)
    x = np.linspace(0.0, 1.0, 33)
    w = E.hopf_lax_primitive(ic, x, t)
    # characteristic foot y: x = y + t u0(y); w = U0(y) + t u0(y)^2 / 2
    y = x.copy()
    for _ in range(60):
        r = y + t * ic(y) - x
        y = y - r / (1 + t * 2 * np.pi * np.cos(2 * np.pi * y))
    wc = ic.antiderivative(y) + t * ic(y) ** 2 / 2
    np.testing.assert_allclose(w, wc, rtol=0, atol=1e-13)


def test_burgers_smooth_order_hopf_lax() -> None:
    ic = E.c1_initial_condition()
    hs, errs = [], []
    for n in [64, 128, 256]:
        grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
        u0 = E.exact_cell_averages(ic, n)
        sch = W.Scheme("rusanov", eps=1e-6)
        u, _, t = W.solve(sch, grid, u0, theta=1.0, tfinal=0.1)
        ue = E.burgers_exact(ic, n, t)
        errs.append(W.norm(grid, u - ue, p=1, weighted=True))
        hs.append(1.0 / n)
    # time error O(dt^3) with dt ~ h dominates
    assert W.estimate_order_of_convergence(np.array(hs), np.array(errs)) >= 2.8


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo"])
def test_burgers_shock_hopf_lax(kind: str) -> None:
    """Past the shock (t = 0.5), against exact Hopf-Lax cell averages."""
    ic = E.c1_initial_condition()
    errs = []
    for n in [64, 128, 256]:
        grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
        u0 = E.exact_cell_averages(ic, n)
        sch = W.Scheme(kind, eps=1e-6)
        u, _, t = W.solve(sch, grid, u0, theta=1.0, tfinal=0.5)
        ue = E.burgers_exact(ic, n, t)
        errs.append(W.rnorm(grid, u, ue, p=1, weighted=True))
    assert errs[-1] < 5.0e-3
    assert errs[0] / errs[-1] > 3.0  # first order through the shock


def test_burgers_flux_sign_error_is_caught() -> None:
    """A planted sign error in the dissipation term must fail the Hopf-Lax
    shock check above (before the shock the term is O(h^5) and nearly
    invisible, so the plant is checked at t = 0.5)."""
    ic = E.c1_initial_condition()
    n = 128
    grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    u0 = E.exact_cell_averages(ic, n)

    def bad_flux(grid, u, a, *, eps, alpha=1.0, rec="wenojs53", flux=W.burgers_flux):
        ul, ur = W.reconstruct_wenojs(u, eps)
        s = np.maximum(np.abs(a[1:]), np.abs(a[:-1]))
        f = 0.5 * (flux(ul)[1:] + flux(ur)[:-1]) + 0.5 * s * (ul[1:] - ur[:-1])
        return np.pad(f, 1)

    orig = W.scalar_flux_rusanov
    try:
        W.scalar_flux_rusanov = bad_flux
        with np.errstate(all="ignore"):
            u, _, t = W.solve(W.Scheme("rusanov", eps=1e-6), grid, u0, theta=1.0,
                              tfinal=0.5)
    finally:
        W.scalar_flux_rusanov = orig
    ue = E.burgers_exact(ic, n, t)
    err = W.rnorm(grid, u, ue, p=1, weighted=True)
    assert not (err < 5.0e-3)


# }}}


# {{{ ES-WENO scheme


def test_es_weno_parameters_closed_form() -> None:
    """``weno.py:265-281`` of a unit state on [a, b]: eps = (b - a) dx^2,
    delta = dx^2."""
    grid = W.make_uniform_cell_grid(-1.0, 2.0, 96, nghosts=2)
    eps, delta = W.es_weno_parameters(grid, np.ones_like(grid.x))
    dx = 3.0 / 96
    assert abs(eps - 3.0 * dx**2) < 1e-15 and abs(delta - dx**2) < 1e-18


def test_esweno_dissipation_closed_forms() -> None:
    """Constant data: no dissipative flux. Linear data: tau = 0, so every
    interior weight is the ideal one and the flux is -delta/8 times the jump
    (away from the zero-padded ends). Random data: the flux opposes the jump
    (its factor sqrt(dw^2 + delta^2)/8 + dw/8 is positive)."""
    delta = 1e-3
    u = np.full(40, 0.7)
    om = W.es_weno_weights(W.WENOJS32, u, eps=1e-6)[0, :]
    np.testing.assert_array_equal(W.esweno_dissipation(u, om, delta), 0.0)

    u = 0.3 + 0.05 * np.arange(40)
    om = W.es_weno_weights(W.WENOJS32, u, eps=1e-6)
    np.testing.assert_allclose(om[:, 1:-1], np.array([[1 / 3], [2 / 3]]) + 0 * om[:, 1:-1],
                               rtol=0, atol=1e-15)
    g = W.esweno_dissipation(u, om[0], delta)
    np.testing.assert_allclose(g[2:-2], -delta / 8 * 0.05, rtol=1e-12, atol=0)
    assert g[0] == 0.0 and g[-1] == 0.0

    u = np.random.default_rng(3).uniform(-1, 1, size=200)
    om = W.es_weno_weights(W.WENOJS32, u, eps=1e-6)[0, :]
    g = W.esweno_dissipation(u, om, delta)[1:-1]
    du = np.diff(u)
    assert np.all(g * du < 0)


def test_esweno_scheme_smooth_order_and_shock() -> None:
    """Burgers with the ES-WENO scheme (bound eps, delta of a unit state)
    against Hopf-Lax before the shock: eps = O(h^2) keeps the weights near
    the ideal ones and delta = h^2 adds O(h^3) dissipation, so the third
    order of the ES-WENO3 faces shows (3.2 measured; a dropped tau, a
    first-order face or an O(h) dissipation would not reach it)."""
    ic = E.c1_initial_condition()
    hs, errs = [], []
    for n in [64, 128, 256]:
        grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=2)
        eps, delta = W.es_weno_parameters(grid, np.ones_like(grid.x))
        u0 = E.exact_cell_averages(ic, n, nghosts=2)
        sch = W.Scheme("esweno32", eps=eps, rec="esweno32", delta=delta)
        u, _, t = W.solve(sch, grid, u0, theta=1.0, tfinal=0.1)
        ue = E.burgers_exact(ic, n, t, nghosts=2)
        errs.append(W.norm(grid, u - ue, p=1, weighted=True))
        hs.append(1.0 / n)
    assert W.estimate_order_of_convergence(np.array(hs), np.array(errs)) >= 2.8


# }}}


def test_esweno_advection_constant_velocity_order() -> None:
    """The ES-WENO advection scheme with a unit velocity (the velocity then
    enters twice, as in ``advection/schemes.py:72``, and f = u) against the
    exact shift of a smooth wave: the Lax-Friedrichs faces of ES-WENO3 with
    eps = O(h^2) and delta = h^2 converge at third order (3.7 measured on
    64-256 cells)."""
    from oracle import exact as E

    hs, errs = [], []
    for n in [64, 128, 256]:
        grid = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=2)
        eps, delta = W.es_weno_parameters(grid, np.ones_like(grid.x))
        vel = np.ones_like(grid.x)
        u0 = W.apply_boundary_periodic(grid, np.sin(2 * np.pi * grid.x))
        sch = W.Scheme("advection-esweno32", eps=eps, velocity=vel, rec="esweno32",
                       delta=delta)
        u, _, t = W.solve(sch, grid, u0, theta=0.5, tfinal=0.25)
        ue = np.sin(2 * np.pi * (grid.x - t))
        errs.append(W.norm(grid, u - ue, p=1, weighted=True))
        hs.append(1.0 / n)
    assert W.estimate_order_of_convergence(np.array(hs), np.array(errs)) >= 2.8


# {{{ Gauss-Legendre cell averages (grid.py:393-435)


@pytest.mark.parametrize("order", [2, 4, 5])
def test_leggauss_cell_average_closed_forms(order: int) -> None:
    """An order-q rule is exact for polynomials of degree 2 q - 1 and not
    for degree 2 q; on sin(2 pi x) it converges at order 2 q to the exact
    average (cos(2 pi x_l) - cos(2 pi x_r)) / (2 pi h). A wrong node
    mapping, a dropped weight or an off-by-one cell fails these."""
    grid = W.make_uniform_cell_grid(-0.25, 1.0, 37, nghosts=3)
    f = grid.f
    h = f[1:] - f[:-1]
    rng = np.random.default_rng(order)

    coeffs = rng.standard_normal(2 * order)  # degree 2 q - 1
    poly = np.polynomial.Polynomial(coeffs)
    prim = poly.integ()
    got = W.leggauss_cell_average(grid, poly, order)
    exact = (prim(f[1:]) - prim(f[:-1])) / h
    np.testing.assert_allclose(got, exact, rtol=1e-12, atol=1e-12)

    # degree 2 q on unit cells, where the rule's error is well above rounding
    wide = W.make_uniform_cell_grid(-1.0, 2.0, 3, nghosts=1)
    fw = wide.f
    top = np.polynomial.Polynomial([0.0] * (2 * order) + [1.0])
    prim = top.integ()
    miss = W.leggauss_cell_average(wide, top, order) - (prim(fw[1:]) - prim(fw[:-1]))
    assert np.min(np.abs(miss)) > 1e-9

    errs = []
    for n in (16, 32):
        g = W.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
        ff = g.f
        exact = (np.cos(2 * np.pi * ff[:-1]) - np.cos(2 * np.pi * ff[1:])) \
            / (2 * np.pi * (ff[1:] - ff[:-1]))
        got = W.leggauss_cell_average(g, lambda x: np.sin(2 * np.pi * x), order)
        errs.append(np.max(np.abs(got - exact)))
    if errs[1] > 1e-14:
        eoc = np.log2(errs[0] / errs[1])
        assert eoc > 2 * order - 0.3, (errs, eoc)


# }}}
