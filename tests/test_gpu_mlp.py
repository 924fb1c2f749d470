"""GPU parity of the learned-weight kernels against the CPU oracle
(``oracle/mlp_np.py`` with the kernels' BF16 rounding points emulated).

The kernels accumulate in FP32 and the oracle in FP64, so a hidden value can
land on the other side of a BF16 rounding boundary; such flips are rare
(about 2^-15 per rounded value) and change a logit by a few 1e-3 at most.
The tolerances below therefore bound the bulk tightly and the outliers
loosely."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_np as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    from pyshocks_amd import _lib

    _lib.load()
    return torch


def _params(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(M.NPARAM) * scale).astype(np.float32)


def _weights(torch, p, tau=20.0):
    from pyshocks_amd.learned import MLPWeights

    return MLPWeights.from_numpy(p, tau)


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 40000])
def test_mlp_weights_vs_oracle(torch_mod, n):
    torch = torch_mod
    from pyshocks_amd import learned

    rng = np.random.default_rng(n)
    p = _params(10)
    x = M.features(rng.standard_normal((n, 5)))
    om, z = learned.weights(_weights(torch, p), torch.tensor(x, device="cuda"), logits=True)
    ref = M.forward(p, x)
    dz = np.abs(z.cpu().numpy() - ref["z"])
    scale = 1.0 + np.abs(ref["z"])
    assert np.max(dz / scale) < 5e-2
    assert np.mean(dz / scale < 1e-5) > 0.99
    w = om.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(w.sum(axis=1), 1.0, atol=1e-6)
    assert np.median(np.abs(w - ref["omega"])) < 1e-6


def test_mlp_layout_single_units(torch_mod):
    """One-hot parameter sets isolate every (hidden unit, input) pair, so a
    wrong lane or slot mapping shows up as a wrong logit, not as noise."""
    torch = torch_mod
    from pyshocks_amd import learned

    x = np.zeros((64, 5), dtype=np.float32)
    for s in range(64):
        x[s, s % 5] = 1.0 if s % 2 == 0 else 0.5
    xt = torch.tensor(x, device="cuda")
    for unit1, unit2 in [(0, 0), (5, 17), (16, 3), (31, 31), (12, 28)]:
        for inp in range(5):
            for out in range(3):
                p = np.zeros(M.NPARAM, np.float32)
                p[M.OFF_W1 + unit1 * 5 + inp] = 1.0
                p[M.OFF_W2 + unit2 * 32 + unit1] = 2.0
                p[M.OFF_W3 + out * 32 + unit2] = 0.75
                _, z = learned.weights(_weights(torch, p), xt, logits=True)
                ref = M.forward(p, x)["z"]
                np.testing.assert_allclose(z.cpu().numpy(), ref, rtol=0, atol=1e-6,
                                           err_msg=f"unit1={unit1} unit2={unit2} in={inp} out={out}")


@pytest.mark.parametrize("n", [100, 5000])
def test_mlp_grad_vs_oracle(torch_mod, n):
    torch = torch_mod
    from pyshocks_amd.learned import Trainer

    rng = np.random.default_rng(20 + n)
    p = _params(11, 0.3)
    rec = np.zeros((n, M.RECORD), dtype=np.float32)
    rec[:, :5] = M.features(rng.standard_normal((n, 5)))
    rec[:, 5:8] = rng.standard_normal((n, 3)) * 0.1
    rec[:, 8:11] = rng.dirichlet(np.ones(3), n)
    rec[:, 11] = rng.uniform(0.0, 1.0, n)
    tr = Trainer(_weights(torch, p), lambda_js=0.05, delta=1e-3)
    g = tr.loss_and_grad(torch.tensor(rec, device="cuda")).cpu().numpy().astype(np.float64)
    loss, gref = M.loss_and_grad(p, rec, lambda_js=0.05, delta=1e-3)
    assert abs(g[-1] - loss) <= 1e-5 * abs(loss)
    assert np.linalg.norm(g[:-1] - gref) <= 2e-3 * np.linalg.norm(gref)
    # every parameter block is reached
    for lo, hi in [(0, 160), (160, 192), (192, 1216), (1216, 1248), (1248, 1344), (1344, 1347)]:
        assert np.linalg.norm(g[lo:hi] - gref[lo:hi]) <= 1e-2 * np.linalg.norm(gref[lo:hi]) + 1e-9


def test_training_reduces_loss(torch_mod):
    """The device training loop fits weights that pick the best candidate
    of synthetic records."""
    torch = torch_mod
    from pyshocks_amd.learned import MLPWeights, Trainer, init_params

    rng = np.random.default_rng(7)
    n = 1 << 15
    x = M.features(rng.standard_normal((n, 5)))
    # candidate k is exact when the largest input sits at position k + 1
    best = np.argmax(x[:, 1:4], axis=1)
    e = np.full((n, 3), 0.3, dtype=np.float32)
    e[np.arange(n), best] = 0.0
    rec = np.zeros((n, M.RECORD), dtype=np.float32)
    rec[:, :5], rec[:, 5:8], rec[:, 8:11], rec[:, 11] = x, e, 1.0 / 3.0, 1.0
    w = MLPWeights.from_numpy(init_params(0))
    tr = Trainer(w, lr=1e-2, lambda_js=0.0, delta=1e-6)
    hist = tr.fit(torch.tensor(rec, device="cuda"), epochs=16, batch=2048)
    assert hist[-1] < 0.25 * hist[0], hist
    loss, _ = M.loss_and_grad(w.params.cpu().numpy(), rec, lambda_js=0.0, delta=1e-6)
    assert abs(loss - hist[-1]) < 0.5 * hist[-1] + 1e-3


def test_training_stops_when_it_diverges(torch_mod):
    """A run whose loss or parameters stop being finite ends with an error
    after that epoch instead of returning weights that turn solutions into
    NaN (here a NaN planted in the initial parameters)."""
    torch = torch_mod
    from pyshocks_amd.learned import MLPWeights, Trainer, init_params

    rng = np.random.default_rng(8)
    n = 4096
    rec = np.zeros((n, M.RECORD), dtype=np.float32)
    rec[:, :5] = M.features(rng.standard_normal((n, 5)))
    rec[:, 5:8] = rng.standard_normal((n, 3)) * 0.1
    rec[:, 8:11] = 1.0 / 3.0
    rec[:, 11] = 1.0
    p = init_params(0)
    p[0] = np.nan
    tr = Trainer(MLPWeights.from_numpy(p), lr=1e-2, lambda_js=0.0, delta=1e-6)
    with pytest.raises(RuntimeError, match="diverged in epoch 1"):
        tr.fit(torch.tensor(rec, device="cuda"), epochs=3, batch=1024)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_learned_reconstruct_vs_oracle(torch_mod, dtype):
    torch = torch_mod
    from pyshocks_amd import learned

    tdt = torch.float64 if dtype == "f64" else torch.float32
    ndt = np.float64 if dtype == "f64" else np.float32
    n = 4096
    x = (np.arange(n + 6) - 3 + 0.5) / n
    u = (0.5 + np.sin(2 * np.pi * x) + 0.4 * np.sin(14 * np.pi * x + 1.0)
         + 0.8 * (x > 0.3) - 0.6 * (x > 0.71)).astype(ndt)
    p = _params(12, 0.3)
    w = _weights(torch, p, tau=20.0)
    ul, ur = learned.learned_reconstruct(w, torch.tensor(u, device="cuda"), 1e-6)
    rl, rr = M.learned_reconstruct(p, 20.0, u, 1e-6)
    v = M.stencils(u.astype(np.float64))
    s = np.ptp(v, axis=1) + 1e-300
    flags = M.shock_flags(M.stencils(u), 1e-6, 20.0)
    assert 4 <= flags.sum() < 40
    scale = np.max(np.abs(u))
    for got, ref in [(ul, rl), (ur, rr)]:
        diff = np.abs(got.cpu().numpy().astype(np.float64) - ref)
        if dtype == "f64":
            # the weights are FP32 (softmax in FP32): errors relative to the
            # local spread of the stencil
            d = diff / s
            assert np.mean(d < 2e-5) > 0.99
        else:
            # FP32 candidates carry rounding relative to |u| itself
            assert np.mean(diff < 2e-6 * scale) > 0.99
            d = diff / (s + 1e-5 * scale)
        assert np.max(d) < 2e-2
        if dtype == "f64":
            np.testing.assert_allclose(got.cpu().numpy()[flags], ref[flags], rtol=1e-12, atol=1e-14)
