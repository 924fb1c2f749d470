"""Pins of the learned-weight oracle (``oracle/mlp_np.py``): finite
differences of the unrounded loss, BF16 rounding, softmax and mirror
invariants, the WENO-JS fallback and the ideal-weight network."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_np as M
from oracle import weno_np


def _records(rng, n, p):
    rec = np.zeros((n, M.RECORD), dtype=np.float32)
    rec[:, :5] = M.features(rng.standard_normal((n, 5)))
    rec[:, 5:8] = rng.standard_normal((n, 3)) * 0.1
    rec[:, 8:11] = M.forward(p, rec[:, :5])["omega"] * 0.5 + 0.5 / 3
    rec[:, 11] = rng.uniform(0.5, 1.0, n)
    return rec


def test_bf16_rounding():
    one = 1.0
    vals = np.array([one, one + 2.0**-8, one + 3 * 2.0**-8, -(one + 3 * 2.0**-8),
                     one + 2.0**-8 + 2.0**-20, 0.0])
    want = np.array([one, one, one + 2.0**-6, -(one + 2.0**-6), one + 2.0**-7, 0.0])
    np.testing.assert_array_equal(M.bf16(vals), want.astype(np.float32))


def test_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = (rng.standard_normal(M.NPARAM) * 0.4).astype(np.float64)
    rec = _records(rng, 64, p)
    kw = dict(lambda_js=0.3, delta=1.0e-2, emulate=False)
    _, g = M.loss_and_grad(p, rec, **kw)
    h = 1.0e-6
    idx = np.concatenate([rng.choice(160, 6, replace=False),            # W1
                          160 + rng.choice(32, 3, replace=False),       # b1
                          192 + rng.choice(1024, 8, replace=False),     # W2
                          1216 + rng.choice(32, 3, replace=False),      # b2
                          1248 + rng.choice(96, 6, replace=False),      # W3
                          [1344, 1345, 1346]])                          # b3
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss_and_grad(p + e, rec, **kw)[0] - M.loss_and_grad(p - e, rec, **kw)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_emulated_gradient_close_to_exact():
    rng = np.random.default_rng(2)
    p = (rng.standard_normal(M.NPARAM) * 0.3).astype(np.float32)
    rec = _records(rng, 512, p)
    _, g0 = M.loss_and_grad(p, rec, lambda_js=0.1, delta=1e-3, emulate=False)
    _, g1 = M.loss_and_grad(p, rec, lambda_js=0.1, delta=1e-3, emulate=True)
    assert np.linalg.norm(g1 - g0) <= 5e-2 * np.linalg.norm(g0)


def test_softmax_and_ideal_network():
    rng = np.random.default_rng(3)
    x = M.features(rng.standard_normal((100, 5)))
    p = (rng.standard_normal(M.NPARAM) * 0.5).astype(np.float32)
    w = M.forward(p, x)["omega"]
    np.testing.assert_allclose(w.sum(axis=1), 1.0, rtol=0, atol=1e-14)
    assert np.all(w >= 0)
    w0 = M.forward(M.ideal_params(), x, emulate=False)["omega"]
    np.testing.assert_allclose(w0, np.broadcast_to(M.IDEAL, w0.shape), rtol=1e-15)


def test_features_mirror_and_scale():
    rng = np.random.default_rng(4)
    v = rng.standard_normal((1000, 5))
    np.testing.assert_array_equal(M.features(v[:, ::-1]), M.features(v)[:, ::-1])
    v32 = v.astype(np.float32)
    np.testing.assert_array_equal(M.features(v32[:, ::-1]), M.features(v32)[:, ::-1])
    x = M.features(v)
    assert np.all(np.max(np.abs(x), axis=1) == 1.0)
    np.testing.assert_allclose(M.features(3.0 * v + 7.0), x, atol=1e-6)
    assert np.all(M.features(np.ones((3, 5))) == 0)


def test_ideal_network_is_linear_interpolant():
    """All-zero parameters: omega = d, so the learned reconstruction is the
    5th-order upwind-biased interpolant [2, -13, 47, 27, -3] / 60."""
    n = 64
    x = np.linspace(0.0, 1.0, n + 6)
    u = np.sin(2 * np.pi * x)
    ul, ur = M.learned_reconstruct(M.ideal_params(), np.inf, u, 1e-6, emulate=False)
    v = M.stencils(u)
    lin_r = v @ (np.array([2.0, -13.0, 47.0, 27.0, -3.0]) / 60.0)
    lin_l = v[:, ::-1] @ (np.array([2.0, -13.0, 47.0, 27.0, -3.0]) / 60.0)
    np.testing.assert_allclose(ur[2:-2], lin_r[2:-2], rtol=0, atol=1e-14)
    np.testing.assert_allclose(ul[2:-2], lin_l[2:-2], rtol=0, atol=1e-14)


def test_flagged_cells_keep_wenojs():
    n = 128
    x = (np.arange(n + 6) - 3 + 0.5) / n
    u = np.where(x < 0.5, 1.0, -0.5) + 0.1 * np.sin(2 * np.pi * x)
    rng = np.random.default_rng(5)
    p = (rng.standard_normal(M.NPARAM) * 0.5).astype(np.float32)
    jl, jr = weno_np.reconstruct_wenojs(u, 1e-6)
    # tau = 0: every cell with any variation is flagged
    ul, ur = M.learned_reconstruct(p, 0.0, u, 1e-6)
    np.testing.assert_array_equal(ul, jl)
    np.testing.assert_array_equal(ur, jr)
    # tau = 20: the cells around the jump are flagged, smooth cells are not
    flags = M.shock_flags(M.stencils(u), 1e-6, 20.0)
    jump = int(np.argmax(x >= 0.5))
    assert flags[jump - 1] and flags[jump]
    assert not np.any(flags[10:jump - 4]) and not np.any(flags[jump + 4:-10])
    ul, ur = M.learned_reconstruct(p, 20.0, u, 1e-6)
    np.testing.assert_array_equal(ur[flags], jr[flags])
    assert np.max(np.abs(ur[~flags][5:-5] - jr[~flags][5:-5])) > 1e-8


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo"])
def test_learned_operator_tau0_is_wenojs_operator(kind):
    """With tau = 0 every varying cell keeps WENO-JS, so the learned operator
    of every flux is the oracle's WENO-JS operator exactly (the flux
    formulas, not just the reconstruction, are the reference's)."""
    n = 64
    grid = weno_np.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    rng = np.random.default_rng(9)
    p = (rng.standard_normal(M.NPARAM) * 0.5).astype(np.float32)
    u = 0.5 + np.sin(2 * np.pi * grid.x) + 0.3 * (grid.x > 0.6) + 0.05 * rng.standard_normal(
        grid.x.size)
    L = M.learned_apply_operator(p, 0.0, kind, grid, u, 1e-6)
    ref = weno_np.apply_operator(weno_np.Scheme(kind, eps=1e-6), grid, 0.0,
                                 weno_np.apply_boundary_periodic(grid, u))
    np.testing.assert_array_equal(L[3:-3], ref[3:-3])


@pytest.mark.parametrize("kind", ["rusanov", "godunov", "lf", "eo"])
def test_learned_operator_conservative(kind):
    n = 64
    grid = weno_np.make_uniform_cell_grid(0.0, 1.0, n, nghosts=3)
    rng = np.random.default_rng(6)
    p = (rng.standard_normal(M.NPARAM) * 0.5).astype(np.float32)
    u = 0.5 + np.sin(2 * np.pi * grid.x) + 0.3 * (grid.x > 0.6)
    L = M.learned_apply_operator(p, 20.0, kind, grid, u, 1e-6)
    assert abs(np.sum(L[grid.i_]) * grid.dx[0]) < 1e-13
