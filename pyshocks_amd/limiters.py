"""Flux limiters of the MUSCL reconstruction.

Mirrors ``limiters.py`` of pyshocks: ``Limiter`` (``:44-72``, with
``is_symmetric``), ``UnlimitedLimiter`` (``:132-152``), ``MINMODLimiter``
(``:166-198``, ``theta`` in ``[1, 2]``), ``MonotonizedCentralLimiter``
(``:215-232``), ``SUPERBEELimiter`` (``:251-268``), ``VanAlbadaLimiter``
(``:284-323``, variants 1 and 2), ``VanLeerLimiter`` (``:345-363``),
``KorenLimiter`` (``:381-389``) and the registry (``:395-436``). The
limiters are evaluated inside the library's MUSCL kernels; ``limiter_code``
gives the ``PSA_LIMITER_*`` code and parameter of one.
"""

from __future__ import annotations

from dataclasses import dataclass, fields
from typing import Any

from pyshocks_amd import _lib


@dataclass(frozen=True)
class Limiter:
    """A flux limiter ``phi(r)`` of the slope ratio ``r``."""

    @property
    def is_symmetric(self) -> bool:
        """Whether ``phi(1 / r) = phi(r) / r``."""
        return False


@dataclass(frozen=True)
class UnlimitedLimiter(Limiter):
    """``phi(r) = 1``."""


@dataclass(frozen=True)
class MINMODLimiter(Limiter):
    """``phi(r) = max(0, min(theta, theta r, (1 + r) / 2))``."""

    theta: float

    def __post_init__(self) -> None:
        if self.theta < 1.0 or self.theta > 2.0:
            raise ValueError(f"'theta' must be in [1, 2]: {self.theta!r}.")

    @property
    def is_symmetric(self) -> bool:
        return True


@dataclass(frozen=True)
class MonotonizedCentralLimiter(Limiter):
    """``phi(r) = max(0, min(2, 2 r, (1 + r) / 2))``."""

    @property
    def is_symmetric(self) -> bool:
        return True


@dataclass(frozen=True)
class SUPERBEELimiter(Limiter):
    """``phi(r) = max(0, min(1, 2 r), min(2, r))``."""

    @property
    def is_symmetric(self) -> bool:
        return True


@dataclass(frozen=True)
class VanAlbadaLimiter(Limiter):
    """``phi(r) = (r^2 + r) / (r^2 + 1)`` (variant 1) or ``2 r / (r^2 + 1)``
    (variant 2), clipped at 0."""

    variant: int

    def __post_init__(self) -> None:
        if self.variant not in {1, 2}:
            raise ValueError(f"Variants 1 and 2 are supported: {self.variant}")

    @property
    def is_symmetric(self) -> bool:
        return self.variant == 1


@dataclass(frozen=True)
class VanLeerLimiter(Limiter):
    """``phi(r) = (r + |r|) / (1 + |r|)``."""

    @property
    def is_symmetric(self) -> bool:
        return True


@dataclass(frozen=True)
class KorenLimiter(Limiter):
    """``phi(r) = max(0, min(2, 2 r, (1 + 2 r) / 3))``."""


def limiter_code(lm: Limiter) -> tuple[int, float]:
    """``(PSA_LIMITER_*, theta)`` of a limiter (theta only for minmod)."""
    if isinstance(lm, MINMODLimiter):
        return _lib.PSA_LIMITER_MINMOD, float(lm.theta)
    if isinstance(lm, MonotonizedCentralLimiter):
        return _lib.PSA_LIMITER_MC, 0.0
    if isinstance(lm, SUPERBEELimiter):
        return _lib.PSA_LIMITER_SUPERBEE, 0.0
    if isinstance(lm, VanAlbadaLimiter):
        return (_lib.PSA_LIMITER_VANALBADA1 if lm.variant == 1
                else _lib.PSA_LIMITER_VANALBADA2), 0.0
    if isinstance(lm, VanLeerLimiter):
        return _lib.PSA_LIMITER_VANLEER, 0.0
    if isinstance(lm, KorenLimiter):
        return _lib.PSA_LIMITER_KOREN, 0.0
    if isinstance(lm, UnlimitedLimiter):
        return _lib.PSA_LIMITER_NONE, 0.0
    raise NotImplementedError(type(lm).__name__)


# {{{ registry

_LIMITERS: dict[str, type[Limiter]] = {
    "default": MINMODLimiter,
    "none": UnlimitedLimiter,
    "minmod": MINMODLimiter,
    "mc": MonotonizedCentralLimiter,
    "superbee": SUPERBEELimiter,
    "vanalbada": VanAlbadaLimiter,
    "vanleer": VanLeerLimiter,
    "koren": KorenLimiter,
}


def limiter_ids() -> tuple[str, ...]:
    return tuple(_LIMITERS.keys())


def make_limiter_from_name(name: str, **kwargs: Any) -> Limiter:
    """Unknown names raise ``ValueError``; fields not in the class are
    ignored (``limiters.py:417-436``)."""
    cls = _LIMITERS.get(name)
    if cls is None:
        raise ValueError(
            f"Flux limiter {name!r} not found. Try one of {', '.join(limiter_ids())}.")

    return cls(**{f.name: kwargs[f.name] for f in fields(cls) if f.name in kwargs})


# }}}
