"""Order-of-convergence bookkeeping (``tools.py:143-262`` of pyshocks)."""

from __future__ import annotations

import numpy as np


def estimate_order_of_convergence(x: np.ndarray, y: np.ndarray) -> tuple[float, float]:
    """Least-squares fit ``y = c x^p``; returns ``(c, p)``."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    assert x.size == y.size
    if x.size <= 1:
        raise RuntimeError("Need at least two values to estimate order.")

    eps = np.finfo(x.dtype).eps
    c = np.polyfit(np.log10(x + eps), np.log10(y + eps), 1)
    return float(10 ** c[-1]), float(c[-2])


class EOCRecorder:
    """Keep track of ``(h, error)`` pairs."""

    def __init__(self, *, name: str = "Error") -> None:
        self.name = name
        self.history: list[tuple[float, float]] = []

    def add_data_point(self, h: float, error: float) -> None:
        self.history.append((float(h), float(error)))

    @property
    def estimated_order(self) -> float:
        if not self.history:
            return float("nan")
        h, e = np.array(self.history).T
        return estimate_order_of_convergence(h, e)[1]

    @property
    def max_error(self) -> float:
        return max((e for _, e in self.history), default=0.0)

    def satisfied(self, order: float, atol: float | None = None, *,
                  slack: float = 0) -> bool:
        if not self.history:
            return True
        if atol is None:
            atol = 1.0e2 * float(np.finfo(np.float64).eps)
        return bool(self.estimated_order >= (order - slack) or self.max_error < atol)

    def __str__(self) -> str:
        lines = [f"{'h':>12} {self.name:>14}"]
        lines += [f"{h:12.5e} {e:14.8e}" for h, e in self.history]
        lines.append(f"order {self.estimated_order:.3f}")
        return "\n".join(lines)
