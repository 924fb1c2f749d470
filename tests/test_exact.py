"""The device Hopf-Lax solver (``pyshocks_amd.exact``, run here on CPU
tensors) against the oracle's independent sampled/golden-section minimiser
(``oracle/exact.py``), and against the closed forms it reduces to."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import exact as E
from pyshocks_amd import exact as X
from pyshocks_amd.funcs import SineStep


def _ics():
    return [
        SineStep(c=0.5, amplitudes=(1.0,), wavenumbers=(1,), phases=(0.0,)),
        SineStep(c=0.3, amplitudes=(0.6, -0.3, 0.2), wavenumbers=(1, 3, 8),
                 phases=(0.4, 2.0, 5.0), step_height=-0.7, step_location=0.35),
        SineStep(c=0.1, amplitudes=(0.4, 0.25), wavenumbers=(2, 5), phases=(1.0, 0.3),
                 step_height=0.8, step_location=0.6),
    ]


def _oracle_ic(ic):
    return E.SineStepIC(c=ic.c, amplitudes=ic.amplitudes, wavenumbers=ic.wavenumbers,
                        phases=ic.phases, step_height=ic.step_height,
                        step_location=ic.step_location)


@pytest.mark.parametrize("t", [0.0, 0.02, 0.2, 0.5])
def test_cell_averages_vs_oracle(t):
    ics = _ics()
    ns = [64, 200, 257]
    got = X.burgers_cell_averages(ics, ns, t, device="cpu")
    for ic, n, g in zip(ics, ns, got):
        ref = E.burgers_exact(_oracle_ic(ic), n, t)
        np.testing.assert_allclose(g.numpy(), ref, rtol=0, atol=2e-11)


def test_zero_time_is_exact_average_and_mass_is_conserved():
    ics = _ics()
    got0 = X.burgers_cell_averages(ics, [128] * 3, 0.0, device="cpu")
    got = X.burgers_cell_averages(ics, [128] * 3, 0.4, device="cpu")
    for ic, a0, a in zip(ics, got0, got):
        np.testing.assert_allclose(a0.numpy(), E.exact_cell_averages(_oracle_ic(ic), 128),
                                   atol=1e-13)
        assert abs(float(a[3:-3].sum() - a0[3:-3].sum()) / 128) < 1e-13


def test_face_values_are_derivative_of_primitive():
    """u(x) = w_x: the point values agree with centred differences of w away
    from shocks, and with the characteristic relation x = y + t u0(y)."""
    ics = _ics()[:2]
    n, t = 512, 0.25
    u = X.burgers_face_values(ics, n, [t, t], device="cpu").numpy()
    for b, ic in enumerate(ics):
        x = np.arange(n + 1) / n
        d = 1e-6
        w1 = E.hopf_lax_primitive(_oracle_ic(ic), x + d, t)
        w0 = E.hopf_lax_primitive(_oracle_ic(ic), x - d, t)
        fd = (w1 - w0) / (2 * d)
        ok = np.abs(fd - u[b]) < 1e-5
        assert np.mean(ok) > 0.97
