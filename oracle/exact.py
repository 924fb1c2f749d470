"""Exact solutions used to pin the oracle independently of any solver.

TEST INFRASTRUCTURE ONLY (imported by ``tests/``, ``bench.py``'s
``cpu_baseline`` leg and the fixture script, never by the product path).

Initial conditions of the family used by the workloads, on the periodic
domain ``[0, 1)``::

    u0(x) = c + sum_m A_m sin(2 pi k_m x + phi_m) + H * 1[frac(x) >= xs]

have a closed-form antiderivative ``U0``, which gives

* exact cell averages ``(U0(x_{i+1/2}) - U0(x_{i-1/2})) / h``;
* the exact solution of linear advection with unit speed, a shift;
* the exact entropy solution of inviscid Burgers through shocks, by the
  Hopf-Lax formula ``w(x, t) = min_y U0(y) + (x - y)^2 / (2 t)`` with
  ``u = w_x``, so that exact cell averages are differences of ``w``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class SineStepIC:
    """``c + sum_m A_m sin(2 pi k_m x + phi_m) + H 1[frac(x) >= xs]``."""

    c: float = 0.0
    amplitudes: tuple[float, ...] = ()
    wavenumbers: tuple[int, ...] = ()
    phases: tuple[float, ...] = ()
    step_height: float = 0.0
    step_location: float = 0.5

    def __post_init__(self) -> None:
        assert len(self.amplitudes) == len(self.wavenumbers) == len(self.phases)
        assert 0.0 < self.step_location < 1.0

    def __call__(self, x: np.ndarray) -> np.ndarray:
        x = np.asarray(x, dtype=np.float64)
        u = np.full_like(x, self.c)
        for a, k, p in zip(self.amplitudes, self.wavenumbers, self.phases):
            u = u + a * np.sin(2.0 * np.pi * k * x + p)
        if self.step_height != 0.0:
            frac = x - np.floor(x)
            u = u + self.step_height * (frac >= self.step_location)
        return u

    def antiderivative(self, y: np.ndarray) -> np.ndarray:
        """``U0(y) = int_0^y u0``, valid on the whole real line."""
        y = np.asarray(y, dtype=np.float64)
        w = self.c * y
        for a, k, p in zip(self.amplitudes, self.wavenumbers, self.phases):
            w = w + a * (np.cos(p) - np.cos(2.0 * np.pi * k * y + p)) / (2.0 * np.pi * k)
        if self.step_height != 0.0:
            fl = np.floor(y)
            frac = y - fl
            xs = self.step_location
            w = w + self.step_height * (fl * (1.0 - xs) + np.maximum(frac - xs, 0.0))
        return w

    def bounds(self) -> tuple[float, float]:
        """Bounds ``umin <= u0 <= umax``."""
        amp = float(np.sum(np.abs(self.amplitudes)))
        lo = self.c - amp + min(self.step_height, 0.0)
        hi = self.c + amp + max(self.step_height, 0.0)
        return lo, hi


def c1_initial_condition() -> SineStepIC:
    """``u0(x) = 0.5 + sin(2 pi x)``."""
    return SineStepIC(c=0.5, amplitudes=(1.0,), wavenumbers=(1,), phases=(0.0,))


def large_initial_condition(seed: int = 0) -> SineStepIC:
    """Sum of 8 sinusoids with amplitudes in [-1, 1], wavenumbers 1..32, random
    phases, plus one step of height in [-1, 1] at a random location."""
    rng = np.random.default_rng(seed)
    amps = rng.uniform(-1.0, 1.0, size=8)
    ks = rng.integers(1, 33, size=8)
    phis = rng.uniform(0.0, 2.0 * np.pi, size=8)
    height = rng.uniform(-1.0, 1.0)
    xs = rng.uniform(0.05, 0.95)
    return SineStepIC(c=0.0, amplitudes=tuple(float(a) for a in amps),
                      wavenumbers=tuple(int(k) for k in ks),
                      phases=tuple(float(p) for p in phis),
                      step_height=float(height), step_location=float(xs))


def interior_faces(a: float, b: float, n: int) -> np.ndarray:
    """The ``n + 1`` faces of the interior cells of a uniform grid."""
    return np.linspace(a, b, n + 1, dtype=np.float64)


def cell_averages_from_primitive(w: np.ndarray, h: float) -> np.ndarray:
    return (w[1:] - w[:-1]) / h


def exact_cell_averages(ic: SineStepIC, n: int, *, a: float = 0.0, b: float = 1.0,
                        nghosts: int = 3) -> np.ndarray:
    """Exact cell averages of ``ic`` on all ``n + 2 g`` cells (ghosts included,
    periodic images)."""
    h = (b - a) / n
    f = a + h * np.arange(-nghosts, n + nghosts + 1, dtype=np.float64)
    return cell_averages_from_primitive(ic.antiderivative(f), h)


def advection_exact(ic: SineStepIC, n: int, t: float, *, velocity: float = 1.0,
                    nghosts: int = 3) -> np.ndarray:
    """Exact cell averages of ``u_t + a u_x = 0`` at time ``t``: a shift."""
    h = 1.0 / n
    f = h * np.arange(-nghosts, n + nghosts + 1, dtype=np.float64)
    return cell_averages_from_primitive(ic.antiderivative(f - velocity * t), h)


def hopf_lax_primitive(ic: SineStepIC, x: np.ndarray, t: float, *,
                       nsamples: int = 4096, ncandidates: int = 4,
                       niterations: int = 100) -> np.ndarray:
    """``w(x, t) = min_y U0(y) + (x - y)^2 / (2 t)`` at every ``x``.

    The minimiser lies in ``[x - t umax, x - t umin]`` (characteristic feet).
    The interval is sampled, every sampled local minimum (up to
    ``ncandidates``, lowest first) is refined by golden-section search in the
    bracket of its two neighbours, and the smallest refined value is kept, so
    that shocks (two competing minima) are resolved to rounding.
    """
    x = np.asarray(x, dtype=np.float64)
    if t == 0.0:
        return ic.antiderivative(x)

    umin, umax = ic.bounds()
    lo = x - t * umax - 1.0e-3
    hi = x - t * umin + 1.0e-3

    s = np.linspace(0.0, 1.0, nsamples)
    y = lo[:, None] + (hi - lo)[:, None] * s[None, :]

    def g(yy: np.ndarray, xx: np.ndarray) -> np.ndarray:
        return ic.antiderivative(yy) + (xx - yy) ** 2 / (2.0 * t)

    gy = g(y, x[:, None])

    # local minima among the interior samples
    is_min = np.zeros_like(gy, dtype=bool)
    is_min[:, 1:-1] = (gy[:, 1:-1] <= gy[:, :-2]) & (gy[:, 1:-1] <= gy[:, 2:])
    masked = np.where(is_min, gy, np.inf)
    order = np.argsort(masked, axis=1)[:, :ncandidates]

    rows = np.arange(x.size)
    best = np.full(x.size, np.inf)
    invphi = (np.sqrt(5.0) - 1.0) / 2.0
    for r in range(order.shape[1]):
        j = order[:, r]
        valid = np.isfinite(masked[rows, j])
        jj = np.clip(j, 1, nsamples - 2)
        a = y[rows, jj - 1]
        b = y[rows, jj + 1]
        c = b - invphi * (b - a)
        d = a + invphi * (b - a)
        gc = g(c, x)
        gd = g(d, x)
        for _ in range(niterations):
            left = gc < gd
            b = np.where(left, d, b)
            a = np.where(left, a, c)
            d_new = np.where(left, c, a + invphi * (b - a))
            c_new = np.where(left, b - invphi * (b - a), d)
            c, d = c_new, d_new
            gc = g(c, x)
            gd = g(d, x)
        val = np.minimum(np.minimum(gc, gd), gy[rows, jj])
        best = np.where(valid, np.minimum(best, val), best)

    # fall back to the sampled minimum where no interior local minimum exists
    best = np.minimum(best, gy.min(axis=1))
    return best


def burgers_exact(ic: SineStepIC, n: int, t: float, *, nghosts: int = 3,
                  **kwargs: int) -> np.ndarray:
    """Exact cell averages of inviscid Burgers at time ``t`` on all
    ``n + 2 g`` cells of a uniform periodic grid on ``[0, 1)``."""
    h = 1.0 / n
    f = h * np.arange(-nghosts, n + nghosts + 1, dtype=np.float64)
    w = hopf_lax_primitive(ic, f, t, **kwargs)
    return cell_averages_from_primitive(w, h)
