"""Host-side pieces of the learned weights that need no GPU: the training
records (torch on the CPU) and the shipped network file."""

from __future__ import annotations

import numpy as np
import torch

from oracle import exact as E
from oracle import mlp_np as M


def test_stencil_records_smooth_and_shock():
    from pyshocks_amd.learned import stencil_records

    n = 400
    ic = E.SineStepIC(c=0.2, amplitudes=(0.5,), wavenumbers=(1,), phases=(0.3,),
                      step_height=0.6, step_location=0.5)
    u = torch.tensor(E.exact_cell_averages(ic, n))[None, :]
    faces = np.arange(n + 1) / n
    uf = torch.tensor(ic(faces))[None, :]
    rec = stencil_records(u, uf, eps=1e-6, tau=20.0).numpy()
    assert rec.shape == (2 * n, 12)
    x = M.features(torch.stack([u[0, k + 1: k + 1 + n] for k in range(5)], dim=-1).numpy())
    np.testing.assert_array_equal(rec[:n, :5], x)
    np.testing.assert_array_equal(rec[n:, :5], x[:, ::-1])
    used = rec[:, 11] > 0
    # the cells around the step are flagged and left out, the rest used
    assert 4 <= np.sum(~used) <= 16
    # smooth right-side records: the ideal weights reproduce the exact face
    # value to fifth order, far better than any single candidate
    r = rec[:n][used[:n]]
    ideal = np.abs(r[:, 5:8] @ np.array([0.1, 0.6, 0.3]))
    assert np.median(ideal / np.max(np.abs(r[:, 5:8]), axis=1)) < 1e-2


def test_shipped_network_file():
    from pyshocks_amd.learned import TRAINED_PATH

    with np.load(TRAINED_PATH, allow_pickle=False) as f:
        p = f["params"]
        assert p.shape == (M.NPARAM,) and p.dtype == np.float32
        assert np.all(np.isfinite(p))
        assert float(f["tau"]) > 0
        # the training loss decreased
        loss = f["loss"]
        assert loss[-1] < loss[0]
    # the network is not the ideal-weight one
    x = M.features(np.random.default_rng(0).standard_normal((256, 5)))
    w = M.forward(p, x)["omega"]
    assert np.max(np.abs(w - M.IDEAL)) > 0.05
